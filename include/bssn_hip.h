/*
 * bssn_hip -- C ABI of the MI355X (gfx950) BSSN right-hand side and time-step library.
 *
 * Every entry point takes device pointers, plain sizes and a HIP stream (hipStream_t passed as
 * void*, 0 for the null stream), enqueues its kernels on that stream and returns 0 on success.
 * A nonzero return is an error code (BSSN_E_*); bssn_last_error() then holds a message for the
 * calling thread. The library never allocates device memory: workspaces are passed in, sized
 * by the *_workspace_bytes queries. Entry points are re-entrant; calls that share buffers on one
 * stream are ordered by that stream.
 *
 * Layouts (the reference's):
 *   state  F[f*P + (x*Sy + y)*Sz + z], P = Sx*Sy*Sz, float64, 24 fields in the order
 *          alpha, chi, K, gt0..5, beta0..2, At0..5, Gt0..2, B0..2
 *          (3D/src/bssn3d/state.py:26-44; index formula gen_fused.py:210-211).
 *   S16    host array of 16 doubles, the reference's scalar pack
 *          (3D/src/bssn3d/fused_rhs_cuda.py:52-55, gen_fused.py:58-60):
 *          eta, C_CAHD, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz.
 *   D      derivative bundle, 138 arrays: d = 3*f + axis (first derivatives of all 24 fields),
 *          d = 72 + 6*s + pair (second derivatives; s over alpha, chi, gt0..5, beta0..2;
 *          pair over 00,01,02,11,12,22; mixed = d1_j(d1_i f)).
 * fd_order is 4, 6 or 8; the Kreiss-Oliger order is fd_order + 2, at most 8 (the reference's
 * pairing and its widest KO stencil, common/src/mcs_common/derivatives.py:45-72): 6, 8, 8.
 */
#ifndef BSSN_HIP_H
#define BSSN_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSSN_HIP_ABI_VERSION 1

#define BSSN_OK 0
#define BSSN_E_ARG 1   /* bad shape, order, pointer or parameter */
#define BSSN_E_HIP 2   /* a HIP runtime error (launch or memory) */

int bssn_abi_version(void);
const char* bssn_last_error(void);

/* Fused right-hand side at every padded point, stencils edge-padded at the array ends.
 * Replaces the reference's XLA FFI handler Fused(stream, F, S) -> OUT
 * (3D/src/bssn3d/cuda/rhs_fused.cu:22-61, bound in fused_rhs_cuda.py:41-72), which is
 * what BSSNSolver(scheme="cuda_fused").rhs calls (3D/src/bssn3d/rhs.py:112-122).
 * ko_sigma = 0 gives the bare RHS (the reference handler); ko_sigma > 0 also adds the
 * Kreiss-Oliger term of BSSNEvolution.rhs (3D/src/bssn3d/evolve.py:86-90). */
int bssn_rhs_f64(const double* F, const double* S16, double* OUT, int Sx, int Sy, int Sz,
                 int fd_order, double ko_sigma, void* stream);

/* bssn_rhs_f64 with the scalar pack in device memory (16 doubles, the S16 order): the
 * reference's handler takes S as a device buffer, Fused(stream, F: (24,Sx,Sy,Sz), S: (16,))
 * (3D/src/bssn3d/cuda/rhs_fused.cu:22-61; S built on the device by _pack_S,
 * fused_rhs_cuda.py:52-55,70-71). Nothing is read on the host, so the call neither copies
 * nor synchronises and can be captured into a graph. The pack's contents are not validated
 * (the host never sees them); results equal bssn_rhs_f64's with the same pack bit for bit. */
int bssn_rhs_sdev_f64(const double* F, const double* S16dev, double* OUT, int Sx, int Sy, int Sz,
                      int fd_order, double ko_sigma, void* stream);

/* Two-pass RHS, pass 1: the 138-derivative bundle at every padded point.
 * Replaces derivative_bundle (3D/src/bssn3d/derivative_bundle.py:31-48) and the device
 * derivative stage deriv_kernel FFI (3D/src/bssn3d/cuda/deriv_2p5d.cu:30-113).
 * h3 = host array {dx, dy, dz}. */
int bssn_deriv_bundle_f64(const double* F, double* D, int Sx, int Sy, int Sz, int fd_order,
                          const double* h3, void* stream);

/* Precision form of pass 1: the same bundle from int8 digits on the INT8 matrix cores
 * (Ozaki scheme I: per 28 x 28 plane tile an offset and a power-of-two scale, `slices` int8
 * digits of 7 bits, banded integer stencils x12 / x144 contracted exactly in INT32 by
 * v_mfma_i32_32x32x32_i8, recombined in FP64). fd_order 4 only; slices 1..7.
 * Replaces the reference's precision-form derivative operator OzakiDerivative
 * (2D/src/mcs2d/schemes/ozaki.py:257-562, 3D/src/mcs3d/schemes/ozaki.py:162-282; there an
 * RNS/CRT form for the MCS solver) applied to the BSSN bundle (derivative_bundle.py:31-48). */
int bssn_deriv_bundle_i8_f64(const double* F, double* D, int Sx, int Sy, int Sz, int fd_order,
                             const double* h3, int slices, void* stream);

/* Two-pass RHS, pass 2: the pointwise algebra over npts points of F (24 x npts) and
 * D (138 x npts). Replaces bssn_rhs_algebra (3D/src/bssn3d/_bssn_rhs_generated.py) and the
 * algebra-only kernel bssn_algebra (3D/src/bssn3d/cuda/gen_algebra_cuda.py:55-96). */
int bssn_algebra_f64(const double* F, const double* D, const double* S16, double* OUT,
                     long npts, void* stream);

/* OUT += Kreiss-Oliger dissipation of all 24 fields, edge-padded (periodic = 0) or wrapped
 * over the interior (periodic = 1, ng ghost cells). Replaces BSSNEvolution._ko
 * (3D/src/bssn3d/evolve.py:76-84). */
int bssn_ko_add_f64(const double* F, double* OUT, int Sx, int Sy, int Sz, int fd_order,
                    double sigma, const double* h3, int periodic, int ng, void* stream);

/* Periodic ghost fill. Replaces BSSNEvolution._periodic_sync (3D/src/bssn3d/evolve.py:65-73). */
int bssn_periodic_sync_f64(double* F, int Sx, int Sy, int Sz, int ng, void* stream);

/* det(gt) = 1, tr(At) = 0, chi/alpha floors over npts points.
 * Replaces BSSNEvolution.enforce (3D/src/bssn3d/evolve.py:93-119). */
int bssn_enforce_f64(double* F, long npts, double chi_floor, double alpha_floor, void* stream);

/* Workspace for one step: integrator 0 = RK4 (3 states), 1 = MSRK rk4_2_2 (2 states). */
size_t bssn_step_workspace_bytes(int integrator, int Sx, int Sy, int Sz);

/* One classic RK4 step of U in place, four fused stage launches, enforcement folded into the
 * last. periodic = 1: every RHS sees periodically wrapped data (the reference's sync before each
 * RHS), interior points are advanced and the ghosts of U are refilled by a periodic sync at the
 * end. periodic = 0: all padded points are advanced with edge-padded stencils and ghosts are not
 * synced (the reference's bc != "periodic"). periodic = 2 ("carried" ghosts): each stage input
 * is synced, every padded point is advanced with edge-padded stencils from it, U keeps its own
 * ghost values as the base of the update and is not synced at the end -- the reference's
 * periodic step value for value, ghosts included (evolve.py:86-90,125-132); its interior equals
 * that of periodic = 1 bit for bit. S16's dt and t entries are replaced by dt and the stage
 * times. Replaces BSSNEvolution.step (3D/src/bssn3d/evolve.py:125-132). */
int bssn_rk4_step_f64(double* U, double* work, int Sx, int Sy, int Sz, const double* S16,
                      int fd_order, double ko_sigma, int ng, int periodic, double t, double dt,
                      double chi_floor, double alpha_floor, void* stream);

/* bssn_rk4_step_f64 with the step time in device memory: each stage reads *tdev (plus its
 * offset) and forms the SSL factor on the device, and *tdev is advanced by dt at the end, so a
 * HIP graph that captured K steps advances the state and the time by K steps per replay.
 * The analogue of the reference's compiled stepping loop with traced t0
 * (BSSNEvolution.evolve, 3D/src/bssn3d/evolve.py:155-194). */
int bssn_rk4_step_devtime_f64(double* U, double* work, int Sx, int Sy, int Sz,
                              const double* S16, int fd_order, double ko_sigma, int ng,
                              int periodic, double* tdev, double dt, double chi_floor,
                              double alpha_floor, void* stream);

/* bssn_rk4_step_f64 with the scalar pack in device memory: dt = S16dev[2], the step time is
 * S16dev[6], the stage times and coefficients are formed on the device, and S16dev[6] is
 * advanced by dt at the end (a captured step replays step after step). The state, the time
 * and every scalar stay on the device: the reference's evolve loop hands Fused a device S per
 * stage (3D/src/bssn3d/evolve.py:86-90,125-132 over rhs_fused.cu:22-61). Results equal
 * bssn_rk4_step_f64's with the same pack, t = S16dev[6], dt = S16dev[2], bit for bit. */
int bssn_rk4_step_sdev_f64(double* U, double* work, int Sx, int Sy, int Sz, double* S16dev,
                           int fd_order, double ko_sigma, int ng, int periodic, double chi_floor,
                           double alpha_floor, void* stream);

/* Radiative (Sommerfeld) outer boundary, for puncture runs on a finite domain: the right-hand
 * side of the ng-deep ghost zones, d_t u = -(x d_x u + y d_y u + z d_z u) / r with the normal
 * derivative one-sided and the tangential ones centred (edge padded), z face over y over x at
 * edges. xmin3 = coordinates of the first interior point (x_i = xmin + (i - ng) h).
 * bssn_sommerfeld_rhs_f64 writes OUT at the ghost points only. bssn_rk4_step_sommerfeld_f64 is
 * bssn_rk4_step_f64 with the interiors advanced by the fused stages (stencils reading the ghost
 * values) and the ghosts by this condition, enforce at every point.
 * The reference has no outer boundary for BSSN (its steps are periodic or leave edge-padded
 * ghosts, 3D/src/bssn3d/evolve.py:65-73,125-132); this follows its MCS solver's
 * bc_sommerfeld (3D/src/mcs3d/main.py:230-250). */
int bssn_sommerfeld_rhs_f64(const double* F, double* OUT, int Sx, int Sy, int Sz, int fd_order,
                            const double* h3, int ng, const double* xmin3, void* stream);
int bssn_rk4_step_sommerfeld_f64(double* U, double* work, int Sx, int Sy, int Sz,
                                 const double* S16, int fd_order, double ko_sigma, int ng,
                                 const double* xmin3, double t, double dt, double chi_floor,
                                 double alpha_floor, void* stream);

/* One MSRK rk4_2_2 step of U in place. fprev holds f(y_{n-1}) on entry and is clobbered;
 * fnext receives f(y_n) (swap the two between steps). periodic as for bssn_rk4_step_f64 (with
 * 2, fprev and fnext hold the ghost rates too). Replaces BSSNEvolution._msrk2_step
 * (3D/src/bssn3d/evolve.py:139-153). */
int bssn_msrk_step_f64(double* U, double* fprev, double* fnext, double* work, int Sx, int Sy,
                       int Sz, const double* S16, int fd_order, double ko_sigma, int ng,
                       int periodic, double t, double dt, double chi_floor, double alpha_floor,
                       void* stream);

/* The stage function f(y, t) of the steppers: RHS + KO, periodic wrap or edge padding.
 * Replaces BSSNEvolution.rhs (3D/src/bssn3d/evolve.py:86-90); used for the MSRK start value. */
int bssn_stage_rhs_f64(const double* Y, double* OUT, int Sx, int Sy, int Sz, const double* S16,
                       int fd_order, double ko_sigma, int ng, int periodic, void* stream);

/* Hamiltonian and momentum constraints at every padded point: HM = 4 arrays (H, M0, M1, M2).
 * Replaces ConstraintSolver.constraints (3D/src/bssn3d/constraints.py:38-51). */
int bssn_constraints_f64(const double* F, double* HM, int Sx, int Sy, int Sz, int fd_order,
                         const double* h3, void* stream);

/* Interior RMS norms {H_l2, M_l2} into the host array out2; synchronises the stream.
 * HM is a 4-array workspace, partial holds bssn_l2_partial_doubles() doubles.
 * Replaces ConstraintSolver.l2 (3D/src/bssn3d/constraints.py:53-62). */
long bssn_l2_partial_doubles(int Sx, int Sy, int Sz, int ng);
int bssn_constraint_l2_f64(const double* F, double* HM, double* partial, int Sx, int Sy, int Sz,
                           int fd_order, const double* h3, int ng, double* out2, void* stream);

/* bssn_constraint_l2_f64 for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): no workspace arguments and no
 * allocation; HM, the partial sums and the totals live in a static device array of
 * bssn_l2_static_capacity_doubles() doubles (10485760, 80 MiB; 4 Sx Sy Sz + 2 (Sx - 2 ng)(Sy - 2 ng) + 2
 * must fit, e.g. 128^3 with ng = 3), which every device the library runs on holds until the
 * process ends. Larger grids are rejected with BSSN_E_ARG. Results equal bssn_constraint_l2_f64's
 * bit for bit. Calls on one device share the array, so calls that overlap on one device race.
 * No counterpart in the reference: its constraint norms allocate through XLA. */
long bssn_l2_static_capacity_doubles(void);
int bssn_constraint_l2_static_f64(const double* F, int Sx, int Sy, int Sz, int fd_order,
                                  const double* h3, int ng, double* out2, void* stream);

/* ------------------------------------------------------------------------- block batches
 * A batch is a set of blocks of different extents and spacings stored in one pool of doubles:
 * block b is a (24, S[3b], S[3b+1], S[3b+2]) state at pool offset off[b], in the single-grid
 * layout above, with spacings h[3b..3b+2]; all blocks share the ghost width ng. One launch
 * evaluates every block. The plan (device memory, caller-owned, bssn_batch_plan_bytes long)
 * holds the block table and the brick lists; the handle (host memory) holds their sizes and the
 * launch setting. A block's results are bit-identical to the single-grid entry points run on
 * that block alone with dx_i = dx, dy, dz set to its spacings.
 * Replaces the slot-batched block-AMR step: fixed-shape slot arrays per level with the RHS
 * vmapped over every slot, active or not (2D/src/mcs2d/amr/evolve.py:659-775,
 * amr/state.py:27-31); here blocks have any extent and only existing blocks are evaluated. */
typedef struct bssn_batch bssn_batch;

/* One ghost point filled from another point of the pool, all 24 fields:
 * pool[dst + f * dst_stride] = pool[src + f * src_stride] (strides = P of each block).
 * Sources must be interior points. */
typedef struct {
  long long dst, src;
  long long dst_stride, src_stride;
} bssn_halo_copy;

/* One point of a node-centred 2:1 prolongation from a parent block (the tensor product, x then
 * y then z, of the 1D refinement fine[2k] = coarse[k], fine[2k+1] = sum_j w_j coarse[k + j],
 * j = -2..3, w = (3, -25, 150, 150, -25, 3)/256, parent reads clamped to its padded block:
 * 3D/src/multipatch/amr/kernels.py:53-108). dst: pool offset of the destination point (field
 * 0); parent: pool offset of the parent block's first padded point (field 0); dst_stride:
 * field stride of the destination block; fine: the index in the 2:1 refinement of the parent's
 * padded block, per axis (0 <= fine < 2 pS); pS: the parent's padded extents (its field stride
 * is pS0 pS1 pS2). */
typedef struct {
  long long dst, parent, dst_stride;
  int fine[3];
  int pS[3];
} bssn_prolong_point;

/* A whole-block prolongation job: the same values as bssn_prolong_point entries for the points
 * of one child block, evaluated separably (x, y and z passes shared by neighbouring points).
 * dst / parent: pool offsets of the child's and the parent's first padded point (field 0);
 * both blocks have padded extent W = BS + 2 ng per axis (W <= 24); corner: the child's octant;
 * skip: bits 0..5 leave out the ghost slab over the interior extent of face -x, +x, -y, +y,
 * -z, +z (those come from neighbours), bit 6 writes the interior too. */
typedef struct {
  long long dst, parent;
  int corner[3];
  int skip;
  int BS, ng;
} bssn_prolong_block;

/* Prolongation points grouped by fine tile: the points of one parent whose fine indices lie in
 * the 4 x 4 x 4 tile starting at origin (multiples of 4 in the parent's 2:1 refinement), with
 * the same values as their bssn_prolong_point entries, bit for bit. The tile's 7^3 coarse
 * window is read once and the x, y and z passes of the tensor-product prolongation are formed
 * in LDS for the whole tile, so neighbouring points share their taps. Points
 * [first, first + count) of the tile point array belong to the tile. */
typedef struct {
  long long parent, first;
  int pS[3];
  int origin[3];
  int count;
  int pad;
} bssn_prolong_tile;

/* One point of a tile: destination (pool offset of field 0), its block's field stride, and its
 * fine index within the tile, (x * 4 + y) * 4 + z. */
typedef struct {
  long long dst;
  int dst_stride;
  int local;
} bssn_prolong_tile_point;

/* A ghost fill program for a block hierarchy: phases run in order; phase p applies the
 * prolongation entries [prolong_end[p-1], prolong_end[p]), the prolongation tiles
 * [tile_end[p-1], tile_end[p]), the block jobs [block_end[p-1], block_end[p]) and the copies
 * [copy_end[p-1], copy_end[p]) (ends cumulative, host arrays of nphase entries; the entry arrays
 * are device memory; any may be empty, and tile_end / tiles / tile_points may be null for a
 * program without tiles). Within a phase the destinations must be disjoint and no entry may
 * read a point the phase writes; a phase may read what earlier phases wrote (a level's
 * prolongation reads its parent level's ghosts). */
typedef struct {
  int nphase;
  const long long* prolong_end;
  const long long* block_end;
  const long long* copy_end;
  const bssn_prolong_point* prolongs;
  const bssn_prolong_block* blocks;
  const bssn_halo_copy* copies;
  const long long* tile_end;
  const bssn_prolong_tile* tiles;
  const bssn_prolong_tile_point* tile_points;
  long long n_tile_points;          /* length of tile_points */
} bssn_ghost_program;

/* Plan size for nblk blocks of padded extents S (3 ints per block); 0 for invalid extents. */
size_t bssn_batch_plan_bytes(int nblk, const int* S, int ng);

/* Build the plan (uploads the block table and brick lists, synchronises the stream) and a
 * handle to it. off: nblk pool offsets; S: 3 * nblk extents (each at most 1024); h: 3 * nblk
 * spacings. Blocks must lie inside the pool and not overlap. */
int bssn_batch_create(bssn_batch** out, void* plan, size_t plan_bytes, int nblk,
                      const long long* off, const int* S, const double* h, int ng,
                      long long pool_doubles, void* stream);
void bssn_batch_destroy(bssn_batch* b);

/* Launch shape of the batched stage kernels: 0 = one workgroup per brick (default); N > 0
 * (a multiple of 8) = N persistent waves pulling bricks from per-XCD queues. */
int bssn_batch_set_launch(bssn_batch* b, int persistent_waves);

/* Interior (interior = 1) or padded point count of the batch; -1 for a null handle. */
long long bssn_batch_points(const bssn_batch* b, int interior);

/* RHS (+ Kreiss-Oliger if ko_sigma > 0) of every block at every padded point, each block
 * edge-padded: bssn_rhs_f64 per block. S16's spacing entries (3, 13..15) are ignored; each
 * block uses its own. */
int bssn_batch_rhs_f64(const bssn_batch* b, const double* F, const double* S16, double* OUT,
                       int fd_order, double ko_sigma, void* stream);

/* Doubles of a batch's derivative bundle: block b's 138 arrays (the single-grid bundle layout
 * of bssn_deriv_bundle_f64 over its padded points) follow block b-1's, each block starting on a
 * 64-double boundary, in the order the blocks were given to bssn_batch_create; -1 for a null
 * handle. */
long long bssn_batch_bundle_doubles(const bssn_batch* b);

/* The INT8-split bundle (bssn_deriv_bundle_i8_f64) of every block of a batch in five launches
 * (one per plane orientation and field group; wave tiles of all blocks in one grid). Each
 * block's arrays equal bssn_deriv_bundle_i8_f64 on that block alone bit for bit. F: the state
 * pool; D: bssn_batch_bundle_doubles(b) doubles. The precision form of the derivatives under
 * the batched-block schedule: the reference's INT8 contraction (2D/src/mcs2d/schemes/
 * ozaki.py:311-377) over its slot-batched AMR step (2D/src/mcs2d/amr/evolve.py:659-775). */
int bssn_batch_deriv_bundle_i8_f64(const bssn_batch* b, const double* F, double* D, int fd_order,
                                   int slices, void* stream);

/* Workspace of a batched step: integrator 0 = RK4 (3 pools). */
size_t bssn_batch_step_workspace_bytes(const bssn_batch* b, int integrator);

/* Each block periodic on its own interior: ghost refill of every block. */
int bssn_batch_periodic_sync_f64(const bssn_batch* b, double* U, void* stream);

/* Apply n halo copies (device array) to a pool of pool_doubles doubles. An entry that would
 * read or write outside the pool is skipped (bssn_count_skipped_entries finds such entries on the
 * host). */
int bssn_batch_halo_f64(double* pool, long long pool_doubles, const bssn_halo_copy* map,
                        long long n, void* stream);

/* How many of n entries (a host array; kind 0 = bssn_halo_copy, 1 = bssn_prolong_point,
 * 2 = bssn_prolong_block, 3 = bssn_prolong_tile, 4 = bssn_prolong_tile_point, 5 = link map
 * entries of bssn_batch_rk4_step_link_f64, 2 ints each) the device
 * kernels would skip on a pool of pool_doubles doubles because they reach outside it (for a
 * block job also: padded extent above 24 or an octant index other than 0 or 1; for a tile: an
 * origin not a multiple of 4 or a count outside 0..4096; for a tile point: a local index outside
 * 0..63); -1 for invalid arguments. The kernels run the same test per entry (a tile's points
 * with its tile). */
long long bssn_count_skipped_entries(int kind, const void* entries, long long n,
                                     long long pool_doubles);

/* One RK4 step of every block of the pool U in place (the reference's per-stage pattern:
 * RHS of all blocks, interior update, ghost sync; 2D/src/mcs2d/amr/evolve.py:745-771).
 * bc = 0: blocks edge-padded (bssn_rk4_step_f64 with periodic = 0, per block);
 * bc = 1: blocks periodic on their own interiors (bssn_rk4_step_f64 with periodic = 1);
 * bc = 2: ghosts are data: interiors advance with stencils reading the ghosts. The stage
 *         inputs start with U's ghosts (held: Dirichlet data from the last sync), then the halo
 *         map (nhalo device entries, may be 0) refills ghosts of every stage input and of U at
 *         the end (the reference's same-level neighbour sync).
 * tdev: if not null, the step time is read from device memory (*tdev, t ignored) and
 * advanced by dt at the end, so a captured step can be replayed step after step. */
int bssn_batch_rk4_step_f64(const bssn_batch* b, double* U, double* work, const double* S16,
                            int fd_order, double ko_sigma, int bc, const bssn_halo_copy* halo,
                            long long nhalo, double t, double dt, double chi_floor,
                            double alpha_floor, double* tdev, void* stream);

/* Block-floating-point storage of the two-pass bundle (the reference's storage-compression
 * form, common/src/mcs_common/bfp_compress.py:28-90 and bfp48.py:7-60, applied to the 138
 * derivative arrays between the passes; all arithmetic stays fp64). Array i is stored as
 * m = clip(rint(d 2^e_i), -2^(k-1), 2^(k-1) - 1) with one exponent per array, k = mant_bits:
 * int32 for k <= 32, three 16-bit two's-complement limbs (low first) for k <= 48. Storage Q is
 * a 1280-byte header (2^-e_i as doubles) followed by the mantissas, array-major like D.
 * bssn_bfp_bytes: Q's size (0 for invalid arguments).
 * bssn_deriv_bundle_absmax_f64: amax[i] = max |d_i| over the padded grid (138 device doubles),
 *   from which the caller forms e_i = floor(log2((2^(k-1) - 1) / amax_i)), 0 for amax_i = 0
 *   (bfp_compress.py:37-46).
 * bssn_deriv_bundle_bfp: pass 1 into BFP storage (exps: host array of 138 exponents).
 * bssn_algebra_bfp_f64: pass 2 reading BFP storage (d = m 2^-e, exact). */
size_t bssn_bfp_bytes(int mant_bits, long long npts);
int bssn_deriv_bundle_absmax_f64(const double* F, double* amax, int Sx, int Sy, int Sz,
                                 int fd_order, const double* h3, void* stream);
int bssn_deriv_bundle_bfp(const double* F, void* Q, const int* exps, int mant_bits, int Sx,
                          int Sy, int Sz, int fd_order, const double* h3, void* stream);
int bssn_algebra_bfp_f64(const double* F, const void* Q, int mant_bits, const double* S16,
                         double* OUT, long long npts, void* stream);

/* Prolongations of n points (device array) into a pool: the reference's prolongate / the halo
 * fill of sync_across_levels (3D/src/multipatch/amr/kernels.py:66-108, sync.py:84-97). */
int bssn_prolong_f64(double* pool, long long pool_doubles, const bssn_prolong_point* pts,
                     long long n, void* stream);

/* n whole-block prolongation jobs (device array). */
int bssn_prolong_blocks_f64(double* pool, long long pool_doubles, const bssn_prolong_block* jobs,
                            long long n, void* stream);

/* Run a ghost fill program on a pool: the hierarchy sync of the reference's N-level step,
 * root level periodic, then each finer level from its parent (prolongation) and from
 * same-level face neighbours (2D/src/mcs2d/amr/evolve.py:719-735,
 * 3D/src/multipatch/amr/sync.py:84-122). */
int bssn_ghost_fill_f64(double* pool, long long pool_doubles, const bssn_ghost_program* prog,
                        void* stream);

/* bssn_batch_rk4_step_f64 with bc = 2 whose ghost refill after every stage and at the end is a
 * ghost fill program: the reference's shared-dt N-level RK4 step over a block hierarchy
 * (2D/src/mcs2d/amr/evolve.py:659-775), with enforcement of the BSSN step folded into the last
 * stage. U must be synced on entry (as every step leaves it), and the program must write every
 * ghost point of every block (the stage buffers' ghosts are not held from U). */
int bssn_batch_rk4_step_prog_f64(const bssn_batch* b, double* U, double* work, const double* S16,
                                 int fd_order, double ko_sigma, const bssn_ghost_program* prog,
                                 double t, double dt, double chi_floor, double alpha_floor,
                                 double* tdev, void* stream);

/* bssn_batch_rk4_step_prog_f64 without the ghost copies between stages: the stage kernels read
 * a block's ghost cells where the link map points -- a same-level neighbour's (or, on the root
 * level, the periodic image's) interior point -- and only the prolongations, with the copies
 * their parent reads need (`linked`, a ghost fill program), refill ghosts after every stage and
 * at the end. `full` (optional) runs once at the end, so that U leaves the step synced exactly
 * as the program step leaves it; interior values equal the program step's bit for bit either
 * way. link: device array of 2 ints per padded point, block b's points at [m_b, m_b + P_b) with
 * m_b = P_0 + ... + P_(b-1) (blocks in the order given to bssn_batch_create,
 * bssn_batch_points(b, 0) points in all): (pool offset of field 0 of the point to read, field
 * stride of its block); read only at ghost points; link_points must equal
 * bssn_batch_points(b, 0), and an entry that would reach outside the pool is not followed (the
 * ghost's own value is read; bssn_count_skipped_entries kind 5 finds such entries). Needs
 * ng >= the stencils' reach and a pool below 2^29 doubles (4 GiB). Same batched step as the reference's slot-batched N-level step
 * (2D/src/mcs2d/amr/evolve.py:659-775); the ghost exchange by reference instead of by copy
 * has no counterpart there. */
int bssn_batch_rk4_step_link_f64(const bssn_batch* b, double* U, double* work,
                                 const double* S16, int fd_order, double ko_sigma,
                                 const int* link, long long link_points,
                                 const bssn_ghost_program* linked,
                                 const bssn_ghost_program* full, double t, double dt,
                                 double chi_floor, double alpha_floor, double* tdev,
                                 void* stream);

/* Refinement indicator of every block of a batch: out[24 b + f] = the largest undivided
 * difference |u(p + e_a) - u(p)| of field f over the block's interior and one ring of ghosts,
 * all three axes (NaN if any is NaN); blocks in the order they were given to
 * bssn_batch_create. Replaces compute_indicator_gradient (3D/src/multipatch/amr/kernels.py:
 * 143-158), per field. Needs ng >= 1. */
int bssn_batch_indicator_f64(const bssn_batch* b, const double* U, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
