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

/** Device rANS stage for every brick; returns per-brick status words. */
function decodeAllRansBricks(dataset) {
  return addon.decodeAllRansBricks(dataset);
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
