'use strict';
// decompose(mesh, options) for Node (src/decompose.ts:407) over the C-ABI library: the
// Node-API addon (cuacd_napi.c) opens libcuacd.so and returns parts shaped like the reference's
// Part (src/part.ts:9-22): {mesh, hull, meshVol, hullVol, hausdorff, stuck, stuckReason?,
// seams?, merged}. Options use the reference's names (src/decompose.ts:26-108), with tree
// settings under options.tree (src/gpu/tree.ts:14-46); options.onProgress receives the
// reference's ProgressInfo (src/decompose.ts:70-82) and options.signal is any object with an
// `aborted` flag (an AbortSignal), read after every progress report: the call then throws an
// Error named AbortError. GpuTree mirrors src/gpu/tree.ts:116-816 (blocking calls), and
// parseOBJ / parseSTL / toOBJ src/mesh.ts:166-242.
const path = require('path');
const addon = require(path.join(__dirname, 'build', 'cuacd_napi.node'));

let opened = false;

// open(libraryPath, symbolPrefix): which library serves the calls (default: the HIP build)
function open(libraryPath, symbolPrefix) {
  addon.open(libraryPath || path.join(__dirname, '..', 'build', 'libcuacd.so'), symbolPrefix || 'cuacd_');
  opened = true;
}

function ready() {
  if (!opened) open();
}

function decompose(mesh, options) {
  ready();
  return addon.decompose(mesh, options || {});
}

class GpuTree {
  // new GpuTree(cfg) (src/gpu/tree.ts:139): cfg with TreeConfig's names, plus useGraph and
  // tuning {hullWgThreads, hullGridGroups, hullGridWaves, tileMin, tileMax, hdTileSamples}
  constructor(cfg) {
    ready();
    this.handle = addon.createTree(cfg || {});
  }
  get planeSlots() { return addon.planeSlots(this.handle); }
  // run(seeds, edgePlanes, {depth, width, nEdgeCuts, readParts, refine}) (:449)
  run(seeds, edgePlanes, opts) { return addon.treeRun(this.handle, seeds, edgePlanes || [], opts || {}); }
  // hullMeshes(meshes, {hullStage, noExport}) (:382)
  hullMeshes(meshes, opts) { return addon.hullMeshes(this.handle, meshes, opts || {}); }
  // hullsWithHausdorff(meshes) (:671)
  hullsWithHausdorff(meshes) { return addon.hullsWithHausdorff(this.handle, meshes); }
  // hausdorff([{a, b}]) (:318)
  hausdorff(pairs) { return addon.hausdorff(this.handle, pairs); }
  destroy() {
    if (this.handle) addon.destroyTree(this.handle);
    this.handle = null;
  }
}

function parseOBJ(text) { ready(); return addon.parseOBJ(text); }
function parseSTL(bytes) { ready(); return addon.parseSTL(bytes); }
function toOBJ(parts) { ready(); return addon.toOBJ(parts); }

module.exports = { open, decompose, GpuTree, parseOBJ, parseSTL, toOBJ };
