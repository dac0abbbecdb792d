'use strict';
// decompose(mesh, options) for Node (src/decompose.ts:407) over the C-ABI library: the
// Node-API addon (cuacd_napi.c) opens libcuacd.so and returns parts shaped like the reference's
// Part (src/part.ts:9-22): {mesh, hull, meshVol, hullVol, hausdorff, stuck, stuckReason?,
// seams?, merged}. Options use the reference's names (src/decompose.ts:26-108), with tree
// settings under options.tree (src/gpu/tree.ts:14-46).
const path = require('path');
const addon = require(path.join(__dirname, 'build', 'cuacd_napi.node'));

let opened = false;

// open(libraryPath, symbolPrefix): which library serves the calls (default: the HIP build)
function open(libraryPath, symbolPrefix) {
  addon.open(libraryPath || path.join(__dirname, '..', 'build', 'libcuacd.so'), symbolPrefix || 'cuacd_');
  opened = true;
}

function decompose(mesh, options) {
  if (!opened) open();
  return addon.decompose(mesh, options || {});
}

module.exports = { open, decompose };
