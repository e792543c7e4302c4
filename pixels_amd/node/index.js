'use strict';
// Node host mirror of the reference's GpuWorld / Simulation API
// (src/gpu/GpuWorld.ts:17-137, src/gpu/Simulation.ts:61-323) over the N-API
// addon pixels_napi.node, which binds the C ABI of include/pixels.h.
// Plain ES2017 so Node 12 runs it without a TypeScript compiler; the types
// are in index.d.ts.

const path = require('path');

const addon = require(path.join(__dirname, 'pixels_napi.node'));

const RAY = { PRIMARY: 0, SHADOW: 1, SHOOT: 2, SEGMENT: 3, MIXED: 4 };
const FACE = { ZERO: 6, NOT_CAST: 0xfe, MISS: 0xff };
// CTRL_NAMES order, src/gpu/layout.ts:48-58
const CTRL_NAMES = [
  'C_FREE_TOP', 'C_IMPACTS', 'C_JOBS', 'C_ALLOCS', 'C_ALLOC_N',
  'C_FLOOD_CUR', 'C_FLOOD_NEXT', 'C_FLOOD_PAR',
  'C_S_STATE', 'C_S_CUR', 'C_S_NEXT', 'C_S_PAR', 'C_NODES', 'C_PENDING',
  'C_S_ITERS', 'C_SEED_N', 'C_CHUNKS', 'C_ARENA', 'C_WRITES', 'C_WRITE_N',
  'C_DIRTY', 'C_PCURSOR', 'C_FLASH', 'C_SPAWNED',
  'S_BILL', 'S_DESTROYED', 'S_ALLOC_FAIL', 'S_FALLING', 'S_SEARCHES', 'S_ISLANDS',
  'S_SEED_DROP', 'S_BUDGET', 'S_HASH_FULL',
];
const CTRL = {};
CTRL_NAMES.forEach((n, i) => { CTRL[n] = i; });

/** Every device resource of one world; the world is uploaded once. */
class GpuWorld {
  /**
   * @param device HIP device ordinal (the reference takes a GPUDevice)
   * @param world  {dx, dy, dz, grid: Uint32Array, pool: Uint16Array, poolUsed, freeBricks?()}
   * @param mats   Float32Array(256*20) in the packMaterials layout
   */
  constructor(device, world, mats, options) {
    const opt = options || {};
    const holes = typeof world.freeBricks === 'function' ? Uint32Array.from(world.freeBricks()) : undefined;
    this.ctx = addon.create({
      dims: [world.dx, world.dy, world.dz],
      grid: world.grid,
      pool: world.pool,
      poolUsed: world.poolUsed,
      holes: holes,
      mats: mats,
      poolCapacity: opt.poolCapacity || 0,
      voxelSize: opt.voxelSize || 1 / 80,
      device: device || 0,
    });
    this.dims = { x: world.dx, y: world.dy, z: world.dz };
    this.bricks = { x: world.dx >> 3, y: world.dy >> 3, z: world.dz >> 3 };
    this.supers = { x: Math.ceil(this.bricks.x / 4), y: Math.ceil(this.bricks.y / 4), z: Math.ceil(this.bricks.z / 4) };
    this.numCells = this.bricks.x * this.bricks.y * this.bricks.z;
    this.initialBricks = world.poolUsed - (holes ? holes.length : 0);
    const used = world.poolUsed;
    this.poolCapacity = opt.poolCapacity || Math.max(used + 98304, used * 2);
    this.uniform = new Float32Array(52);
  }

  /** Uniform block (shared.wgsl:5-19), written once per frame. */
  writeUniforms(u) { addon.setUniforms(this.ctx, u || this.uniform); }

  /** Renderer light list (render_common.wgsl:8-9). */
  writeLights(lights) { addon.setLights(this.ctx, lights); }

  destroy() {
    if (this.ctx) addon.destroy(this.ctx);
    this.ctx = null;
  }
}

/** Dispatch / queue API of Simulation.ts over the library. */
class Simulation {
  constructor(gw) {
    this.gw = gw;
    this.stats = { bill: 0, destroyed: 0, falling: 0, particles: 0, liveBricks: 0, allocFailures: 0, searches: 0,
      islands: 0, searching: false };
    this.onStats = null;
    this.pendingImpacts = 0;
  }

  async init() { /* kernels are compiled into the library: nothing to build at run time */ }

  /** Simulation.probe (Simulation.ts:161-196): decoded voxels of a box, x fastest. */
  async probe(origin, size) { return addon.probe(this.gw.ctx, origin, size); }

  /** pushImpact (sim.wgsl:203-211): Float32Array of 16-float records. */
  pushImpacts(impacts) {
    addon.pushImpacts(this.gw.ctx, impacts);
    this.pendingImpacts += impacts.length / 16;
  }

  /** The hot-path stages of one frame: the edit chain run to quiescence. */
  encode() {
    addon.islandPass(this.gw.ctx);
    this.pendingImpacts = 0;
  }

  /** Unsupported components of the last pass: {keys, counts}, sorted by key. */
  islandLabels() { return addon.islandLabels(this.gw.ctx); }

  /** Ray queries: rays is an ArrayBuffer of 32-byte records; returns 16-byte hits. */
  trace(kind, rays) { return addon.trace(this.gw.ctx, kind, rays); }

  /** afterSubmit (Simulation.ts:289-322): counters -> stats. */
  afterSubmit() {
    const r = addon.readCtrl(this.gw.ctx);
    const u = r.ctrl;
    this.stats = {
      bill: u[CTRL.S_BILL] / 10,
      destroyed: u[CTRL.S_DESTROYED],
      falling: u[CTRL.S_FALLING],
      particles: r.args[37],
      liveBricks: this.gw.poolCapacity - u[CTRL.C_FREE_TOP],
      allocFailures: u[CTRL.S_ALLOC_FAIL],
      searches: u[CTRL.S_SEARCHES],
      islands: u[CTRL.S_ISLANDS],
      searching: u[CTRL.C_S_STATE] !== 0,
    };
    if (this.onStats) this.onStats(this.stats);
    return this.stats;
  }
}

module.exports = { GpuWorld, Simulation, RAY, FACE, CTRL, CTRL_NAMES, addon };
