'use strict';
// Drives the Node.js module the way web/app.js drives the reference
// (load -> decodeAllRansBricks -> decode -> render) and writes the results
// as raw little-endian files for the Python parity test to compare.
const fs = require('fs');
const path = require('path');
const fcs = require(path.join(__dirname, '..', '..', 'fast-compressed-segmentation_amd', 'node'));

const [csdbPath, outDir, brickSize] = process.argv.slice(2);
const buf = fs.readFileSync(csdbPath);
const ds = fcs.loadDatasetFromArrayBuffer(new Uint8Array(buf.buffer, buf.byteOffset, buf.length), Number(brickSize));
const h = ds.header;

async function main() {
// app.js:75-82: decodeAllRansBricks resolves to one nibble stream per brick
const nibbles = await fcs.decodeAllRansBricks(ds);
const st = ds.ransStatus;
fs.writeFileSync(path.join(outDir, 'rans_status.bin'), Buffer.from(st.buffer));
fs.writeFileSync(path.join(outDir, 'nibbles.bin'), Buffer.concat(nibbles.map((a) => Buffer.from(a.buffer))));
fs.writeFileSync(path.join(outDir, 'nibble_lengths.bin'),
                 Buffer.from(Uint32Array.from(nibbles.map((a) => a.length)).buffer));
const nibOk = nibbles.every((a) => a instanceof Uint8Array);
const one = fcs.readNibbles(ds, 7);
const readNibblesMatches = one.length === nibbles[7].length && one.every((x, i) => x === nibbles[7][i]);
// every brick at every LOD through readBrick
const parts = [];
for (let b = 0; b < h.numBricks; b++) {
  for (let t = 0; t <= h.levels; t++) parts.push(Buffer.from(fcs.readBrick(ds, b, t).buffer));
}
fs.writeFileSync(path.join(outDir, 'bricks.bin'), Buffer.concat(parts));
// both stages again in one call, then every brick once more
const vst = fcs.decodeVolume(ds);
fs.writeFileSync(path.join(outDir, 'volume_status.bin'), Buffer.from(vst.buffer));
const parts2 = [];
for (let b = 0; b < h.numBricks; b++) {
  for (let t = 0; t <= h.levels; t++) parts2.push(Buffer.from(fcs.readBrick(ds, b, t).buffer));
}
fs.writeFileSync(path.join(outDir, 'bricks_after_volume.bin'), Buffer.concat(parts2));
const g = h.gridSize;
const cam = { position: [g / 2, g / 2, -1], forward: [0, 0, 1], right: [1, 0, 0], up: [0, -1, 0],
              tanHalfFov: 0, orthoHalfHeight: g / 2 };
const lods = new Uint8Array(h.numBricks).fill(h.levels);
const f = fcs.render(ds, cam, 256, 256, lods);
fs.writeFileSync(path.join(outDir, 'label.bin'), Buffer.from(f.label.buffer));
fs.writeFileSync(path.join(outDir, 'voxel.bin'), Buffer.from(f.voxel.buffer));
// shaded frame, then every label hidden (a transparent frame)
const shaded = fcs.render(ds, cam, 256, 256, lods, true);
fs.writeFileSync(path.join(outDir, 'rgba.bin'), Buffer.from(shaded.rgba.buffer));
fcs.updateLightingOptions(ds, { lightAngle: 25 });
const labels = fcs.getDatasetLabels(ds);
fcs.applyLabelVisibility(ds, labels);
fcs.applyLabelHover(ds, labels[0]);
const hiddenAll = fcs.render(ds, cam, 256, 256, lods, true).rgba.every((x) => x === 0);
fcs.applyLabelVisibility(ds, []);
fcs.applyLabelHover(ds, null);
// forceFullBrickRefresh / reDecodeCachedBricks: the resident set decoded again
const before = fcs.cacheResident(ds);
const probe = fcs.readBrick(ds, 5, h.levels);
const refreshed = fcs.forceFullBrickRefresh(ds);
const again = fcs.reDecodeCachedBricks(ds);
const after = fcs.cacheResident(ds);
const probe2 = fcs.readBrick(ds, 5, h.levels);
const sameLabels = refreshed === again && after.bricks.length === before.bricks.length &&
                   probe.every((x, i) => x === probe2[i]);
let threw = false;
try { fcs.loadDatasetFromArrayBuffer(new Uint8Array(64)); } catch (e) { threw = /Invalid file format/.test(e.message); }
console.log(JSON.stringify({ header: h, badMagicThrows: threw, marchMs: f.marchMs, numLabels: labels.length,
                              hiddenAllTransparent: hiddenAll, nibbleArrays: nibbles.length,
                              nibblesAreUint8Arrays: nibOk, readNibblesMatches,
                              residentBefore: before.bricks.length, refreshResident: refreshed,
                              refreshSameLabels: sameLabels }));
}

main().catch((e) => { console.error(e); process.exit(1); });
