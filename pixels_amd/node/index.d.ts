// Types of the Node host mirror (index.js) of src/gpu/GpuWorld.ts and
// src/gpu/Simulation.ts over the C ABI in include/pixels.h.

export interface HostWorld {
  dx: number;
  dy: number;
  dz: number;
  grid: Uint32Array;
  pool: Uint16Array;
  poolUsed: number;
  freeBricks?(): number[];
}

export interface WorldOptions {
  poolCapacity?: number;
  voxelSize?: number;
}

export declare class GpuWorld {
  constructor(device: number, world: HostWorld, mats: Float32Array, options?: WorldOptions);
  /** Scene rasterised on the GPU from 168-byte px_prim records (shapes.ts:334-414). */
  static fromPrims(device: number, dims: [number, number, number], prims: ArrayBuffer | Uint8Array,
                   mats: Float32Array, options?: WorldOptions): GpuWorld;
  readonly dims: { x: number; y: number; z: number };
  readonly bricks: { x: number; y: number; z: number };
  readonly supers: { x: number; y: number; z: number };
  readonly numCells: number;
  readonly poolCapacity: number;
  readonly initialBricks: number;
  readonly uniform: Float32Array;
  writeUniforms(u?: Float32Array): void;
  writeLights(lights: Float32Array): void;
  destroy(): void;
}

export interface SimStats {
  bill: number;
  destroyed: number;
  falling: number;
  particles: number;
  liveBricks: number;
  allocFailures: number;
  searches: number;
  islands: number;
  searching: boolean;
}

export declare class Simulation {
  constructor(gw: GpuWorld);
  stats: SimStats;
  onStats: ((s: SimStats) => void) | null;
  init(): Promise<void>;
  probe(origin: [number, number, number], size: [number, number, number]): Promise<Uint32Array>;
  pushImpacts(impacts: Float32Array): void;
  /** Without a mask the edit chain to quiescence; with STAGE bits those stages in order. */
  encode(stages?: number): void;
  particlesInit(maxParticles?: number): void;
  spawn(records: ArrayBuffer | Uint8Array): void;
  readParticles(first?: number, count?: number): Uint8Array;
  readAlive(): Uint32Array;
  pushWrites(pairs: Uint32Array): void;
  readChunks(arenaCap?: number): { table: Uint32Array; arena: Uint32Array };
  islandLabels(): { keys: Uint32Array; counts: Uint32Array };
  trace(kind: number, rays: ArrayBuffer | Uint8Array): Uint8Array;
  afterSubmit(): SimStats;
}

export declare class Renderer {
  constructor(gw: GpuWorld);
  /** fsMain colour (RGBA floats) of every 32-byte ray record. */
  render(rays: ArrayBuffer | Uint8Array): Float32Array;
  readDensity(): Uint8Array;
}

export declare const STAGE: { EDIT: 1; PARTICLES: 2; FRAME_BEGIN: 4; WRITES: 8; RESOLVE: 16; CHUNKS: 32 };
export declare const STAGE_ALL: 63;
export declare const RAY: { PRIMARY: 0; SHADOW: 1; SHOOT: 2; SEGMENT: 3; MIXED: 4 };
export declare const FACE: { ZERO: 6; NOT_CAST: 0xfe; MISS: 0xff };
export declare const CTRL: Record<string, number>;
export declare const CTRL_NAMES: string[];
