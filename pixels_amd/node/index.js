'use strict';
// Node host mirror of the reference's GpuWorld / Simulation API
// (src/gpu/GpuWorld.ts:17-137, src/gpu/Simulation.ts:61-323) over the N-API
// addon pixels_napi.node, which binds the C ABI of include/pixels.h.
// Plain ES2017 so Node 12 runs it without a TypeScript compiler; the types
// are in index.d.ts.

const path = require('path');

const addon = require(path.join(__dirname, 'pixels_napi.node'));

const RAY = { PRIMARY: 0, SHADOW: 1, SHOOT: 2, SEGMENT: 3, MIXED: 4 };
// stage bits of Simulation.encode (include/pixels.h PX_STAGE_*), run in the
// reference's order: frame begin, edit, resolve, chunks, particles, writes
const STAGE = { EDIT: 1, PARTICLES: 2, FRAME_BEGIN: 4, WRITES: 8, RESOLVE: 16, CHUNKS: 32 };
const STAGE_ALL = 63;
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

  /**
   * The scene rasterised on the GPU (apply() of shapes.ts:334-414 for every
   * primitive in order): prims is an ArrayBuffer of 168-byte px_prim records.
   */
  static fromPrims(device, dims, prims, mats, options) {
    const opt = options || {};
    const gw = Object.create(GpuWorld.prototype);
    gw.ctx = addon.createFromPrims({ dims: dims, mats: mats, poolCapacity: opt.poolCapacity || 0,
      voxelSize: opt.voxelSize || 1 / 80, device: device || 0 }, prims);
    gw.dims = { x: dims[0], y: dims[1], z: dims[2] };
    gw.bricks = { x: dims[0] >> 3, y: dims[1] >> 3, z: dims[2] >> 3 };
    gw.supers = { x: Math.ceil(gw.bricks.x / 4), y: Math.ceil(gw.bricks.y / 4), z: Math.ceil(gw.bricks.z / 4) };
    gw.numCells = gw.bricks.x * gw.bricks.y * gw.bricks.z;
    gw.initialBricks = 0;
    gw.poolCapacity = opt.poolCapacity || 0;  // 0: sized by the library from the bricks the scene needs
    gw.uniform = new Float32Array(52);
    return gw;
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

  /**
   * The stages of one frame (Simulation.encode, Simulation.ts:200-273).
   * Without a mask: the edit chain run to quiescence; with a mask of STAGE
   * bits: those stages in the reference's order.
   */
  encode(stages) {
    if (stages === undefined) addon.islandPass(this.gw.ctx);
    else addon.simStep(this.gw.ctx, stages);
    this.pendingImpacts = 0;
  }

  /** Debris particle ring of maxParticles slots (0 = MAX_PARTICLES, config.ts:38). */
  particlesInit(maxParticles) {
    addon.particlesInit(this.gw.ctx, maxParticles || 0);
    this.maxParticles = maxParticles || 262144;
  }

  /** spawn() (sim.wgsl:166-170) of 32-byte particle records, in order. */
  spawn(records) { addon.spawn(this.gw.ctx, records); }

  /** Ring slots [first, first + count) as 32-byte records. */
  readParticles(first, count) { return addon.readParticles(this.gw.ctx, first || 0, count || this.maxParticles); }

  /** The frame's alive list (the particle draw list). */
  readAlive() { return addon.readAlive(this.gw.ctx, this.maxParticles || 262144); }

  /** pushWrite (sim.wgsl:186-192) of (key, value) pairs. */
  pushWrites(pairs) { addon.pushWrites(this.gw.ctx, pairs); }

  /** Chunk table (16 words per chunk) and arena (key, value | chunk << 16). */
  readChunks(arenaCap) { return addon.readChunks(this.gw.ctx, arenaCap); }

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

/** fsMain (raymarch.wgsl:387-450) over explicit camera rays. */
class Renderer {
  constructor(gw) { this.gw = gw; }

  /** RGBA float colour of every 32-byte ray record (debug views by U.flags.w). */
  render(rays) { return addon.render(this.gw.ctx, rays); }

  /** The AO density texture, one byte per brick cell. */
  readDensity() { return addon.readDensity(this.gw.ctx, this.gw.numCells); }
}

module.exports = { GpuWorld, Simulation, Renderer, RAY, FACE, STAGE, STAGE_ALL, CTRL, CTRL_NAMES, addon };
