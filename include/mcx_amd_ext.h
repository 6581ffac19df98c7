/*
 * mcx_amd_ext.h -- extensions of the MI355X engine beyond the reference ABI.
 *
 * None of these has a counterpart in the reference; they are library-wide
 * switches and statistics that the reference's Config cannot carry without
 * changing its layout.  Plain C, no HIP types.
 */
#ifndef MCX_AMD_EXT_H
#define MCX_AMD_EXT_H

#include <stddef.h>

#include "mcx_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fluence tally precision used by subsequent mcx_run_simulation calls:
 *  0  FP32 atomics with the reference's MAX_ACCUM shadow half (reference
 *     semantics, src/mcx_core.cu:2246-2253)
 *  1  FP32 hi + FP32 lo compensated (TwoSum) tally, merged in FP64
 *  2  FP64 no-return atomics (default: same 8 bytes per voxel-gate as the
 *     two FP32 halves, the most accurate, and the fastest because no
 *     deposit waits for an atomic's return value)
 *  3  FP64, one private copy per XCD updated with workgroup-scope atomics
 *     (completed in that XCD's L2), the copies summed in FP64 at merge
 * Returns the previous mode, or -1 for an invalid mode. */
int  mcx_amd_set_tally_mode(int mode);

/* Collect per-run work counters (steps, deposits, scatters, launches) in
 * subsequent runs; costs one atomic per wave at lane exit. */
void mcx_amd_set_counters(int enable);

/* Hot-cell box (default on): in subsequent mcx_run_simulation calls with
 * tally mode 1, 2 or 3, the tally cells of the first gate of each time
 * window in a box of voxels at the source are summed in FP64 in each
 * workgroup's LDS and added to the tally when the workgroup ends.  These
 * cells take the most deposits per cell (about one per photon at a pencil
 * source's voxel), and global atomics on one address serialise, so
 * without the box they bound the kernel.  The box is as large as the LDS
 * left at full occupancy allows.  Returns the previous setting. */
int  mcx_amd_set_hot_box(int enable);

/* Blocked FP64 tally: in subsequent mcx_run_simulation calls with tally
 * mode 2, the tally is stored in 2 x 2 x 2 voxel blocks (the 8 cells of a
 * block in one 64-byte line), each lane sums its deposits into the block
 * it is in, in LDS, and when a lane moves to another block the wave adds
 * the blocks left by its lanes to the tally with 8 lanes to a block; the
 * device's atomic unit takes one line per request.  The output is the same
 * volume (merged back to the reference's order).  mode 0 off, 1 on, 2
 * (default) on for label volumes whose mean scattering coefficient times
 * the voxel length is at most 2 (more scattering per voxel means fewer
 * deposits per step, where the blocks only add work).  Returns the
 * previous mode, or -1 for an invalid one. */
int  mcx_amd_set_tally_blocks(int mode);

/* K independent runs of one forward simulation in one launch of the
 * single-simulation kernel (no counterpart in the reference; for callers
 * that need several independent estimates of one volume, such as a
 * denoiser's inputs).  K is a power of two from 1 to 64; virtual thread
 * idx belongs to run idx & (K - 1), so every run has its own RNG streams
 * and nthread / K of them, and cfg->nphoton photons are launched per run.
 * Each run keeps its own FP64 tally and launched energy and is normalised
 * like mcx_run_simulation's flux; the outputs are reduced on the device to
 * time-integrated maps.  Forward runs only (no detected photons, replay,
 * respin or gate groups; output type flux).  cw: K x nvox floats (x
 * fastest), host or device memory; energy: K pairs (escaped, launched);
 * stats: [0] tally clear
 * ms, [1] transport kernel ms, [2] reduction ms, [3] kernel launch+sync
 * wall ms.  Returns 0, or the error id (message via mcx_last_error). */
int  mcx_amd_run_replicas(Config* cfg, GPUInfo* gpu, int k, float* cw, double* energy, double* stats);

/* Lane refill under autopilot: mcx_run_simulation and
 * mcx_amd_run_replicas with autopilot launch one lane per resident SIMD
 * slot and R times as many virtual threads (nthread = R x lanes, each with
 * its own RNG stream and the reference's photon split); a lane whose
 * virtual thread has launched all its photons takes the next unstarted
 * one, so lanes keep working until the last virtual threads start rather
 * than idling until the slowest lane of their wave is done.  The results
 * are those of a launch with one lane per virtual thread and the same
 * nthread.  mode 0: off (R = 1); 1 (default): R = photons per lane / 64,
 * at most 8, so every virtual thread keeps at least 64 photons; 2: R =
 * photons per lane, at most 32 (about one photon per virtual thread, the
 * shortest tail).  The first random number of every stream the host
 * seeding makes is biased (mean 0.4914; the reference draws the same),
 * which shifts results by O(1/photons per thread) (cube60's absorbed
 * fraction: +0.018/P relative, below 2.8e-4 at P >= 64): mode 2 trades
 * that for speed on small runs.  Returns the previous mode, or -1 for a
 * mode outside 0-2. */
int  mcx_amd_set_refill(int mode);

/* The per-thread seed words of the host seeding (src/mcx_core.cu:3247-3251,
 * :3381-3387): srand(seed), then for each word (rand() << 16) |
 * (rand() << 1) | (rand() >> 14), evaluated left to right, as
 * mcx_run_simulation draws them for each launch.  ncalls consecutive draws
 * of counts[i] words each from one generator are written back to back to
 * out (large draws are generated on several host threads).  Host only; no
 * counterpart in the reference, which calls rand() inline.  Returns 0. */
int  mcx_amd_seed_words(unsigned int seed, unsigned int* out, const unsigned long long* counts, int ncalls);

/* Numbering of replayed photons (seed = SEED_FROM_FILE) in subsequent
 * mcx_run_simulation calls.  0 (default): photon k of virtual thread idx is
 * saved photon idx*threadphoton + min(idx, oddphotons) + k, which visits
 * every saved photon once.  1: the reference's numbering
 * idx*threadphoton + min(idx, oddphotons - 1) + k (src/mcx_core.cu:1206,
 * :1701, :2021): identical while threads outnumber photons; otherwise one
 * photon is replayed twice and the last one never, and a run in which every
 * thread takes the same count is rejected (the reference reads the entry
 * before the first saved photon).  Returns the previous setting. */
int  mcx_amd_set_replay_index(int reference);

/* Statistics of the last mcx_run_simulation call on this process:
 *  out[0] kernel time in ms (HIP events, sum over launches)
 *  out[1] launch+sync wall time in ms (the reference's runtime definition)
 *  out[2] steps  out[3] deposits  out[4] scatters  out[5] launches
 *  out[6] vector registers per lane  out[7] lanes launched
 *  out[8] workgroup size  out[9] kernel launches
 *  out[10] lanes that hit the "should never happen" exit
 *  out[11] tally cells in the hot-cell box (0: off)
 *  out[12] dynamic LDS bytes per workgroup
 *  out[13] deposits summed in the hot-cell box (of out[3]; with counters)
 *  out[14] 1 when the tally was blocked (mcx_amd_set_tally_blocks)
 *  out[15] global tally atomic requests of the lanes (with counters; a
 *          blocked tally's block flush counts once)
 *  out[16] step-loop iterations summed over waves (with counters): out[2]
 *          / (64 x out[16]) is the mean fraction of a wave's lanes
 *          stepping
 *  out[17] lanes launched (out[7] virtual threads; more with refill)
 * Returns the number of values written (<= n). */
int  mcx_amd_last_stats(double* out, int n);

/* Scattered-atomic rate of the device (the roof of the fluence tally):
 * nblocks x 256 lanes each add 1 at `nper` uniformly random indices of an
 * `nelem`-element array.  mode 0: FP64 agent scope, no return; 1: FP32 no
 * return; 2: FP32 returning; 3: FP64 workgroup scope into one private copy
 * per XCD; coalesced: 4 FP64 with 8 lanes per random aligned 64-byte
 * block, 5 FP64 with a wave's 64 lanes on 64 consecutive elements, 6 FP32
 * no return with 16 lanes per 64-byte block (nelem >= 64).  out[0]
 * atomics/s, out[1] ms, out[2] array sum / atomics issued (1 = no lost
 * update), out[3..10] workgroups per XCC id.  Returns 0 on success. */
int  mcx_amd_atomic_bench(int mode, size_t nelem, unsigned int nper, unsigned int nblocks, double* out);

/* Options and results of mcx_run_batch.  Inputs: zero selects the default
 * (tally 1; photons_per_lane and block_size autotuned; one execution).
 * Outputs are written back. */
typedef struct McxBatchOptions {
    int   tally_mode;        /* 1: FP32 hi/lo compensated, merged in FP64 (default); 2: FP64 */
    int   photons_per_lane;  /* virtual threads per instance = ceil(nphoton / this), rounded up to 64;
                                0: each Config's own nthread when autopilot is off, else autotuned */
    int   block_size;        /* workgroup size, multiple of 64; 0: autotuned */
    int   use_graph;         /* capture clear + transport + energy sums + merge in a hipGraph */
    int   repeat;            /* executions of the sequence (results are the last one's); <= 1: once */
    int   download;          /* 1: copy each normalised fluence (nvox x gates) to cfg->exportfield;
                                2: copy the time-integrated map instead, tstep x the sum over gates,
                                   reduced on the device (nvox floats);
                                0: none.  exportfield is allocated when NULL (refused when the
                                buffers would exceed 80 % of the available host memory) */
    float tune_fraction;     /* photon fraction used by the autotuner's trial runs (0: 1/16) */
    int   chosen_photons_per_lane;  /* out */
    int   chosen_block_size;        /* out */
    int   vgprs;                    /* out: vector registers per lane of the batched kernel */
    float kernel_ms;         /* out: transport kernel time of the last execution (graph: whole graph) */
    float total_ms;          /* out: last execution, clear to merge (and, with download 2, the
                                time-integrated reduction) */
    double photons;          /* out: photons simulated per execution */
    double lanes;            /* out: virtual threads over all instances */
    double counters[4];      /* out, with mcx_amd_set_counters(1): steps, deposits, scatters, launches of the
                                last execution; zeros otherwise */
    int   tally_cache_cells; /* per-wave LDS tally cache (hot-cell box at each instance's source): < 0 none,
                                > 0 at most this many cells, 0 autotuned (the largest that fits when the
                                autotuner does not run) */
    int   chosen_tally_cache_cells; /* out: cells per wave used (0: none) */
    double wave_iterations;  /* out, with counters: step-loop iterations summed over waves; counters[0] /
                                (64 x this) is the mean fraction of a wave's lanes stepping */
} McxBatchOptions;

/* Runs `ncfg` independent simulations in one persistent kernel launch.
 * Extends mcx_run_simulation (src/mcx_core.h:40-42; same preconditions on
 * each Config, e.g. after mcx_validatecfg) to many instances: each gets its
 * own volume, media, source, time gates, photon count and seed, and on
 * return its energytot/energyesc/energyabs/normalizer/runtime and, with
 * `download`, its normalised fluence (or time-integrated map) in
 * exportfield (overwritten, not accumulated).  An instance with autopilot off and the same nthread as a
 * single-simulation run reproduces that run's per-thread trajectories.
 * Not batched: detected photons, respin != 1, labels >= 128.
 * `gpu` may be NULL (device from cfgs[0]->deviceid).  Returns 0, or a
 * negative error code with the message in mcx_last_error(). */
int  mcx_run_batch(Config** cfgs, int ncfg, GPUInfo* gpu, McxBatchOptions* opt);

/* sizeof/offsetof values of the ABI structs as compiled into the library:
 * sizeof(Config), sizeof(GPUInfo), sizeof(History), then the offsets of
 * Config.srcpos, .session, .flog, .his, .exportfield, .srcparam1,
 * .deviceid, .energytot, .bc, .srcdata, then sizeof(McxBatchOptions) and
 * offsetof(McxBatchOptions, photons).  Returns the count available. */
int  mcx_amd_abi_layout(size_t* out, int n);

#ifdef __cplusplus
}
#endif

#endif
