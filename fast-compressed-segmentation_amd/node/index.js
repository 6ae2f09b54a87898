'use strict';
/**
 * Node.js module over libfcs (HIP decoder + DDA marcher on MI355X).
 *
 * Mirrors the entry points web/app.js imports (web/app.js:9-12):
 *   loadDatasetFromArrayBuffer(arrayBuffer)  (web/scripts/datasetLoader.js:39-145)
 *   decodeAllRansBricks(dataset)             (web/scripts/decode.js:17-26)
 * and replaces the WebGPU decode/march passes (web/scripts/gpuPasses.js:15-51)
 * with requestBricks / decodeAll / render on the device.
 */
const path = require('path');

const addon = require(path.join(__dirname, 'fcs.node'));

let defaultCtx = null;
function context() {
  if (!defaultCtx) defaultCtx = addon.createContext(0);
  return defaultCtx;
}

/** Parse a .csdb ArrayBuffer / Uint8Array and upload it; throws
 * "Invalid file format" on a bad magic like the reference loader. */
function loadDatasetFromArrayBuffer(buffer, brickSizeOverride = 0, ctx = null) {
  const bytes = buffer instanceof Uint8Array ? buffer : new Uint8Array(buffer);
  return addon.loadCsdb(ctx || context(), bytes, brickSizeOverride);
}

/** decodeAllRansBricks (web/scripts/decode.js:17-26): resolves to one
 * Uint8Array per brick (brick order), its rANS-decoded nibble stream with two
 * nibbles per byte, the first in the high half -- what web/app.js:75-82
 * stores as each brick's encodedData. The rANS stage runs on the device and
 * its output stays resident there for decoding; the arrays are host copies.
 * Per-brick status words of the stage: decodeRansStatus / dataset.ransStatus.
 * Rejects with "rANS decode failed" when a brick's stream is corrupt. */
async function decodeAllRansBricks(dataset) {
  const status = decodeRansStatus(dataset);
  const bad = status.findIndex((s) => s !== 0);
  if (bad >= 0) throw new Error(`rANS decode failed (brick ${bad}, status 0x${status[bad].toString(16)})`);
  return addon.allNibbles(dataset);
}

/** The device rANS stage for every brick (fcs_decode_rans): per-brick
 * status words (0 = ok), also kept as dataset.ransStatus. */
function decodeRansStatus(dataset) {
  const status = addon.decodeRans(dataset);
  dataset.ransStatus = status;
  return status;
}

/** One brick's nibble stream (an element of decodeAllRansBricks' result). */
function readNibbles(dataset, brick) {
  return addon.readNibbles(dataset, brick);
}

/** Decode again every brick that is resident, at the LOD it holds (its
 * finest level for a pyramid), after dropping the whole cache. Returns the
 * number of bricks decoded. */
function redecodeResident(dataset) {
  const res = addon.cacheResident(dataset);
  addon.cacheClear(dataset);
  if (res.bricks.length) addon.requestBricks(dataset, res.bricks, res.lods);
  return res.bricks.length;
}

/** forceFullBrickRefresh (web/scripts/renderLoop.js:502-559): reset the
 * brick cache and immediately decode again the bricks that were resident.
 * The reference needs this after a visibility change because its cache holds
 * colours; this cache holds labels and colours are applied while marching,
 * so the images are already current -- the call keeps the reference's
 * effect on the cache. */
function forceFullBrickRefresh(dataset) {
  return redecodeResident(dataset);
}

/** reDecodeCachedBricks (web/scripts/renderLoop.js:565-595): decode every
 * resident brick again. */
function reDecodeCachedBricks(dataset) {
  return redecodeResident(dataset);
}

// Per-dataset label visibility state: applyLabelVisibility and
// applyLabelHover each keep the other's setting, like the reference
// (web/scripts/webgpuSetup.js:867-882).
const visibility = new WeakMap();
function visState(dataset) {
  if (!visibility.has(dataset)) visibility.set(dataset, { hidden: new BigUint64Array(0), highlight: undefined });
  return visibility.get(dataset);
}

/** updateLightingOptions (web/scripts/webgpuSetup.js:186-222): lightAngle in
 * degrees, ambient, diffuseStrength; unspecified fields keep their values. */
function updateLightingOptions(dataset, options = {}) {
  const cur = Object.assign({ lightAngle: 25, ambient: 0.5, diffuseStrength: 0.8 }, dataset.__lighting || {}, options);
  dataset.__lighting = cur;
  addon.setLighting(dataset, cur.lightAngle, cur.ambient, cur.diffuseStrength);
}

/** applyLabelVisibility (web/scripts/webgpuSetup.js:867-875): hidden labels
 * become transparent in shaded frames. */
function applyLabelVisibility(dataset, hiddenLabelList) {
  const st = visState(dataset);
  st.hidden = BigUint64Array.from(hiddenLabelList.map((l) => BigInt(l)));
  addon.setLabelVisibility(dataset, st.hidden, st.highlight);
  return { applied: true, hiddenCount: st.hidden.length };
}

/** applyLabelHover (web/scripts/webgpuSetup.js:877-882): brighten one label (null: none). */
function applyLabelHover(dataset, labelOrNull) {
  const st = visState(dataset);
  st.highlight = labelOrNull === null || labelOrNull === undefined ? undefined : BigInt(labelOrNull);
  addon.setLabelVisibility(dataset, st.hidden, st.highlight);
}

/** getDatasetLabels (web/scripts/webgpuSetup.js:57-59): the transfer
 * function's labels, ascending. */
function getDatasetLabels(dataset) {
  return Array.from(addon.transferFunction(dataset).labels);
}

module.exports = {
  deviceCount: addon.deviceCount,
  createContext: addon.createContext,
  loadDatasetFromArrayBuffer,
  decodeAllRansBricks,
  decodeRansStatus,
  readNibbles,
  forceFullBrickRefresh,
  reDecodeCachedBricks,
  cacheResident: addon.cacheResident,
  cacheClear: addon.cacheClear,
  requestBricks: addon.requestBricks,
  decodeAll: addon.decodeAll,
  decodeVolume: addon.decodeVolume,
  readBrick: addon.readBrick,
  orbitCamera: addon.orbitCamera,
  render: addon.render,
  updateLightingOptions,
  applyLabelVisibility,
  applyLabelHover,
  getDatasetLabels,
  transferFunction: addon.transferFunction,
};
