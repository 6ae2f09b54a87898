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

module.exports = {
  deviceCount: addon.deviceCount,
  createContext: addon.createContext,
  loadDatasetFromArrayBuffer,
  decodeAllRansBricks,
  requestBricks: addon.requestBricks,
  decodeAll: addon.decodeAll,
  readBrick: addon.readBrick,
  orbitCamera: addon.orbitCamera,
  render: addon.render,
};
