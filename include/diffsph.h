/*
 * diffsph.h -- C ABI of the MI355X (gfx950) particle hot path.
 *
 * Every entry point takes plain device pointers, sizes and scalars, launches
 * HIP kernels on the given stream (a hipStream_t passed as void*, NULL = the
 * null stream) and returns a dsph_status.  No entry point allocates device
 * memory: scratch comes from a caller-provided workspace whose size is given
 * by the matching *_workspace() query.  Entry points are stateless and
 * thread-safe; calls on one stream are ordered, calls on different streams
 * that share output buffers race.
 *
 * Indexing: particle counts are int64 in the ABI, indices on the device are
 * int32 (so a single set is limited to 2^31-1 particles and 2^31-1 pairs).
 * Positions and velocities are [n,2] row-major float32 (x0,y0,x1,y1,...),
 * scalars are [n] float32.
 *
 * Each declaration cites the reference interface it replaces (paths are
 * relative to the diffSPH repository root).
 */
#ifndef DIFFSPH_AMD_H
#define DIFFSPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dsph_status {
  DSPH_OK = 0,
  DSPH_ERR_ARG = 1,        /* invalid argument (null pointer, bad size, bad mode) */
  DSPH_ERR_DIM = 2,        /* unsupported dimension / grid shape */
  DSPH_ERR_LAUNCH = 3,     /* a HIP call or kernel launch failed */
  DSPH_ERR_NOT_CONVERGED = 4,
  DSPH_ERR_CAPACITY = 5,   /* output buffer or workspace too small */
  DSPH_ERR_DEVICE = 6      /* no usable HIP device */
} dsph_status;

/* Support-radius mode of a neighbor search, as the reference's supportMode
 * (neighborhood.h, countNeighborsForParticle): pair kept when
 * SCATTER: |x_ij| / h_j <= 1, GATHER: |x_ij| / h_i <= 1,
 * SYMMETRIC: |x_ij| / ((h_i+h_j)/2) <= 1. */
enum { DSPH_MODE_SCATTER = 0, DSPH_MODE_GATHER = 1, DSPH_MODE_SYMMETRIC = 2 };

/* Axis-aligned 2-D domain.  Periodic axes wrap positions with
 * torch.remainder semantics (v2/modules/neighborhood.py:407-430) and pair
 * offsets with v2/math.py:4-6 `mod`. */
typedef struct dsph_domain {
  float min[2];
  float max[2];
  int32_t periodic[2];
} dsph_domain;

/* Dense cell grid used by the neighbor search (replaces the compact hash
 * map of neighborDebug.ipynb cell 4 / torchCompactRadius.buildDataStructure,
 * v2/modules/neighborhood.py:403).  Cells are squares of side `hcell`
 * (>= the largest search radius); cell (cx,cy) has linear key cy*ncell[0]+cx. */
typedef struct dsph_grid {
  int32_t ncell[2];
  float hcell;
  int32_t pad_;
} dsph_grid;

/* A neighbor list plus what is needed to evaluate every pair on the fly.
 * Two layouts of the same rows:
 *  - CSR (slice_off == NULL): row i holds nbr[offsets[i] .. offsets[i+1]);
 *  - sliced ELL, 64 rows per slice (slice_off != NULL): entry n of row i is
 *    nbr[slice_off[i/64] + 64*n + i%64] for n < counts[i]; a slice is padded
 *    to its longest row.  One slice is one wave64: every load of a
 *    neighbor index (or of a per-pair value stored in the same layout) is a
 *    contiguous 256-byte access.
 * Entries of a row are in the same order in both layouts.  Positions are
 * the *wrapped* positions the search used. */
typedef struct dsph_pairs {
  int64_t nq;               /* query particles (rows)          */
  int64_t nr;               /* reference particles             */
  const int32_t* offsets;   /* [nq+1] CSR row offsets          */
  const int32_t* nbr;       /* neighbor indices (either layout) */
  const float* qpos;        /* [nq,2] wrapped query positions  */
  const float* rpos;        /* [nr,2] wrapped reference pos.   */
  const float* qh;          /* [nq] or NULL -> h_uniform       */
  const float* rh;          /* [nr] or NULL -> h_uniform       */
  float h_uniform;
  int32_t mode;             /* DSPH_MODE_*                     */
  dsph_domain domain;
  const int32_t* counts;    /* [nq] row lengths (sliced layout) */
  const int32_t* slice_off; /* [ceil(nq/64)+1] or NULL (CSR)    */
  /* Optional 16-bit copy of nbr for the DFSPH pressure sweeps (NULL: not
   * used): entry k holds d with j = (i + d) mod nr, |d| <= 32767, written
   * by dsph_neighbor_compact16; nbr16_overflow is its device flag (nonzero:
   * a delta did not fit and the sweeps read nbr). */
  const int16_t* nbr16;
  const int32_t* nbr16_overflow;
  /* Optional neighbor tiles for the DFSPH pressure sweeps (tile_slot NULL:
   * not used; nq == nr), written by dsph_neighbor_tiles.  The neighbors of
   * the rows of workgroup b (rows 256 b .. 256 b + 255) lie in R disjoint
   * runs of consecutive particles, tile_runs[b * (2 DSPH_TILE_MAX_RUNS + 1)
   * ...] = {R, start_0, length_0, start_1, ...}, tile_n[b] particles in
   * all; entry k of such a row holds in tile_slot[k] the position of its
   * neighbor in the runs laid end to end.  The sweeps stage a workgroup's
   * runs of neighbor records in LDS once and read them there.  tile_n[b] < 0
   * or > tile_cap: that workgroup's rows read their neighbors from global
   * memory (the nbr16 / nbr entries). */
  const uint16_t* tile_slot;
  const int32_t* tile_runs;
  const int32_t* tile_n;
  int32_t tile_cap;         /* 1 .. DSPH_TILE_MAX_CAP */
  int32_t tile_pad_;
} dsph_pairs;

#define DSPH_TILE_ROWS 256
#define DSPH_TILE_MAX_CAP 4096
#define DSPH_TILE_MAX_RUNS 16

/* Particles per grid cell beyond which the state is taken as diverged or
 * collapsed (NaN positions land in cell 0): such a cell is left unsorted
 * and the neighbor search skips candidate ranges longer than 3 cells of
 * this size, so that every kernel's work stays bounded.  The fixed-width
 * fills report it with *row_overflow = 2 (1: a row longer than width). */
#define DSPH_CELL_MAX 4096

/* 16-bit neighbor deltas for the pressure sweeps (halves their index
 * stream): out[k] = (nbr[k] - i) reduced mod nr into [-nr/2, nr/2], for
 * every entry of every row in the list's own layout (padding entries are
 * not written; `capacity` is the length of `out`).  *overflow (device
 * int32) is cleared, then set when some |d| > `limit` (at most 32767) or
 * an entry lies beyond `capacity`; the sweeps then keep reading nbr.  Sorted in cell
 * order, a neighbor of i lies within about (cells per grid row + 1) x
 * (particles per cell) slots of i, or as far across the periodic seam, so
 * the deltas fit for grids up to ~2000 cells per row at ~16 particles per
 * cell.  No counterpart in the reference (its lists are int64 pairs). */
int dsph_neighbor_compact16(const dsph_pairs* pairs, int32_t limit,
                            int16_t* out, int64_t capacity,
                            int32_t* overflow, void* stream);

/* Neighbor tiles of a square list (nq == nr) for the DFSPH pressure sweeps
 * (the tile_* fields of dsph_pairs), from the cell grid the list was
 * searched on (`grid`, `cell_start` of dsph_grid_build): for each
 * workgroup b of 256 rows, the particles of the cells around its rows'
 * cells (the 3 x 3 stencil of the search, periodic as the domain is) as
 * sorted disjoint runs in tile_runs[b * (2 DSPH_TILE_MAX_RUNS + 1) ...],
 * their count in tile_n[b], and for every entry k of those rows the
 * position of its neighbor in the runs in slot[k] (`slot_capacity` >= the
 * layout's length).  A workgroup whose runs hold more than `cap` particles,
 * whose rows span more than 4 rows of cells, whose runs number more than
 * DSPH_TILE_MAX_RUNS, or one of whose neighbors lies outside its runs gets
 * tile_n[b] = count or -1 and is not tiled: the sweeps read its neighbors
 * from global memory.  Neighbor indices are read from nbr16 when it is
 * valid, else from nbr.  cap: 1 .. DSPH_TILE_MAX_CAP.  No counterpart in
 * the reference: a layout of the neighbor list that the DFSPH sweeps
 * (v2/modules/dfsph.py:35-50) read. */
int dsph_neighbor_tiles(const dsph_pairs* pairs, const dsph_grid* grid,
                        const int32_t* cell_start, int32_t cap, uint16_t* slot,
                        int64_t slot_capacity, int32_t* tile_runs,
                        int32_t* tile_n, void* stream);

/* Sliced-ELL layout from row counts: slice_off[s] = 64 * sum over earlier
 * slices of their longest row; slice_off[nslices] = padded entry count,
 * also copied to *total_host when not NULL (synchronises the stream).
 * With total_host, a padded entry count beyond INT32_MAX returns
 * DSPH_ERR_DIM (the int32 slice offsets would wrap); without it the caller
 * guarantees the count fits. */
size_t dsph_sell_workspace(int64_t nq);
int dsph_sell_layout(const int32_t* counts, int64_t nq, int32_t* slice_off,
                     int64_t* total_host, void* workspace,
                     size_t workspace_bytes, void* stream);

/* Copy a sliced list into CSR order: out_nbr[offsets[i] + n] = entry n of
 * row i. */
int dsph_sell_to_csr(const dsph_pairs* pairs, int32_t* out_nbr, void* stream);

const char* dsph_status_string(int status);
/* 4: dsph_pairs carries the neighbor-tile fields. */
int dsph_abi_version(void);

/* ---------------------------------------------------------------------- */
/* Neighbor search                                                          */
/* ---------------------------------------------------------------------- */

/* Host-only: grid for a domain and a largest search radius `hsearch`.
 * Periodic axes get floor(L/hsearch) cells of side L/ncell (reference
 * compute_h, neighborhoodDebug.ipynb cell 2: one side for both axes, the
 * larger of the two); a periodic axis needs >= 3 cells. */
int dsph_grid_dims(const dsph_domain* domain, float hsearch, dsph_grid* grid);

size_t dsph_grid_build_workspace(int64_t n, const dsph_grid* grid);

/* Sort a particle set into cell order (replaces sortReferenceParticles +
 * buildCompactHashMap, neighborhoodDebug.ipynb cells 3-4).  Outputs:
 *   cell_start[ncell+1]  first sorted index of every cell,
 *   sort_index[n]        original index of sorted slot p (stable: ascending
 *                        original index inside a cell),
 *   rank[n]              sorted slot of original particle i (may be NULL),
 *   sorted_pos[n,2]      wrapped positions in sorted order,
 *   sorted_h[n]          supports in sorted order (NULL if h is NULL). */
int dsph_grid_build(const dsph_domain* domain, const dsph_grid* grid,
                    const float* pos, const float* h, int64_t n,
                    int32_t* cell_start, int32_t* sort_index, int32_t* rank,
                    float* sorted_pos, float* sorted_h,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Count, for each query particle, the reference particles within its support
 * (replaces countNeighbors, neighborhood.cpp:65-227, and the counting pass
 * of computeNeighborhood, v2/cppSrc/computeNeighborhood.h:115-146).
 * `qpos_wrapped` != 0 when qpos already holds wrapped positions.
 * The predicate is evaluated in float32 as
 *   d = mod(wrap(x_i) - wrap(x_j)); r = sqrt_rn(d.x*d.x + d.y*d.y) (no FMA);
 *   keep iff r /_rn h_mode <= 1
 * (v2/modules/neighborhood.py:445-453). */
int dsph_neighbor_count(const dsph_domain* domain, const dsph_grid* grid,
                        const int32_t* cell_start, const float* sorted_pos,
                        const float* sorted_h, int64_t nref,
                        const float* qpos, const float* qh, int64_t nq,
                        int32_t qpos_wrapped, float h_uniform, int32_t mode,
                        int32_t* counts, void* stream);

/* Fill the neighbor list: row i receives the sorted reference indices of
 * its neighbors, cells visited dy-major (-1,0,+1) then dx, particles in
 * sorted order (replaces buildNeighborList, neighborhood.cpp:253-435).
 * With slice_off == NULL the list is CSR (offsets[nq+1] = exclusive scan of
 * the counts); otherwise it is the sliced-ELL layout of dsph_sell_layout
 * (offsets unused, padding entries set to -1).  The list must hold
 * `capacity` entries; a layout needing more is rejected with
 * DSPH_ERR_CAPACITY before anything is written. */
int dsph_neighbor_fill(const dsph_domain* domain, const dsph_grid* grid,
                       const int32_t* cell_start, const float* sorted_pos,
                       const float* sorted_h, int64_t nref,
                       const float* qpos, const float* qh, int64_t nq,
                       int32_t qpos_wrapped, float h_uniform, int32_t mode,
                       const int32_t* offsets, const int32_t* slice_off,
                       int64_t layout_total, int32_t* nbr, int64_t capacity,
                       void* stream);

/* One-pass neighbor list in a fixed-width sliced layout (no count pass, no
 * host read of the list size): slice s starts at s * 64 * width (slice_off
 * [ceil(nq/64)+1] is written), entry n of row i at slice_off[i/64] + 64 n +
 * i%64, counts[i] = the row's entry count (capped at width).  When a row
 * has more than `width` entries *row_overflow is set and the list is
 * incomplete (build the two-pass list instead).  Entries past a row's count
 * are not written.  With nbr16 the 16-bit deltas of
 * dsph_neighbor_compact16 are written in the same pass (rows = reference
 * slots: nq == nref and qpos_wrapped != 0, else DSPH_ERR_ARG;
 * *delta_overflow as there).  Both flags are cleared first.
 * capacity >= ceil(nq/64) * 64 * width.  No counterpart in the reference. */
int dsph_neighbor_fill_fixed(const dsph_domain* domain, const dsph_grid* grid,
                             const int32_t* cell_start, const float* sorted_pos,
                             const float* sorted_h, int64_t nref,
                             const float* qpos, const float* qh, int64_t nq,
                             int32_t qpos_wrapped, float h_uniform, int32_t mode,
                             int32_t width, int32_t* counts, int32_t* slice_off,
                             int32_t* nbr, int64_t capacity, int16_t* nbr16,
                             int32_t* delta_overflow, int32_t* row_overflow,
                             void* stream);

/* dsph_neighbor_fill_fixed whose entries are map[j] for the reference slot
 * j (map: the sort index of the reference grid, nref entries), i.e. the
 * list dsph_remap_neighbors would make of it, written in the same pass.  No
 * 16-bit deltas.  No counterpart in the reference. */
int dsph_neighbor_fill_fixed_map(const dsph_domain* domain, const dsph_grid* grid,
                                 const int32_t* cell_start, const float* sorted_pos,
                                 const float* sorted_h, int64_t nref,
                                 const float* qpos, const float* qh, int64_t nq,
                                 int32_t qpos_wrapped, float h_uniform, int32_t mode,
                                 int32_t width, int32_t* counts, int32_t* slice_off,
                                 int32_t* nbr, int64_t capacity, const int32_t* map,
                                 int32_t* row_overflow, void* stream);

/* dsph_neighbor_fill_fixed that also builds the list's neighbor tiles (the
 * tile_* fields of dsph_pairs; the list must be square: nq == nref and
 * qpos_wrapped): the run tables tile_runs / tile_n first (from the cells
 * alone), then, in the fill itself, each entry's 16-bit tile slot beside
 * it (slot: slot_capacity >= ceil(nq / 64) * 64 * width entries).  The
 * same tables and slots as dsph_neighbor_tiles(cap = tile_cap) makes of
 * the filled list, without its pass over the entries; a tile of more than
 * 4095 particles is left unusable (tile_n = -1).  nbr16 (optional) as in
 * dsph_neighbor_fill_fixed.  No counterpart in the reference. */
int dsph_neighbor_fill_fixed_tiles(const dsph_domain* domain, const dsph_grid* grid,
                                   const int32_t* cell_start, const float* sorted_pos,
                                   const float* sorted_h, int64_t nref,
                                   const float* qpos, const float* qh, int64_t nq,
                                   int32_t qpos_wrapped, float h_uniform, int32_t mode,
                                   int32_t width, int32_t* counts, int32_t* slice_off,
                                   int32_t* nbr, int64_t capacity, int16_t* nbr16,
                                   int32_t* delta_overflow, int32_t tile_cap,
                                   uint16_t* slot, int64_t slot_capacity,
                                   int32_t* tile_runs, int32_t* tile_n,
                                   int32_t* row_overflow, void* stream);

/* dsph_neighbor_fill that also writes the 16-bit modular deltas of
 * dsph_neighbor_compact16 (d = j - i reduced mod nref; |d| <= 32767, else
 * *overflow is set) in the same pass, for a list whose rows are the sorted
 * reference slots themselves (a set searched against its own grid and not
 * remapped afterwards).  *overflow is cleared first.  Requires nq == nref and
 * qpos_wrapped != 0 (the queries are the grid's own sorted positions),
 * else DSPH_ERR_ARG.  No counterpart in the reference. */
int dsph_neighbor_fill16(const dsph_domain* domain, const dsph_grid* grid,
                         const int32_t* cell_start, const float* sorted_pos,
                         const float* sorted_h, int64_t nref,
                         const float* qpos, const float* qh, int64_t nq,
                         int32_t qpos_wrapped, float h_uniform, int32_t mode,
                         const int32_t* offsets, const int32_t* slice_off,
                         int64_t layout_total, int32_t* nbr, int16_t* nbr16,
                         int64_t capacity, int32_t* overflow, void* stream);

/* out = getPeriodicPosition(pos) (v2/modules/neighborhood.py:407-430):
 * periodic axes x' = remainder(x - min, max - min) + min in float32. */
int dsph_wrap_positions(const dsph_domain* domain, const float* pos, int64_t n,
                        float* out, void* stream);

size_t dsph_scan_workspace(int64_t n);
/* out[0..n] = exclusive prefix sum of in[0..n) (out[n] = total).  If
 * total_host != NULL the total is also copied to host memory (this
 * synchronises the stream, as the reference's `.item()` at
 * v2/cppSrc/computeNeighborhood.cpp:58 does).  Partial sums are carried in
 * int64: with total_host, a total beyond INT32_MAX (or a negative input)
 * returns DSPH_ERR_DIM instead of wrapped offsets; without it the caller
 * guarantees the total fits (as the fixed-width list does by its size
 * check). */
int dsph_scan_exclusive(const int32_t* in, int64_t n, int32_t* out,
                        int64_t* total_host, void* workspace,
                        size_t workspace_bytes, void* stream);

/* Gather helpers for moving per-particle data between original and sorted
 * order: dst[p] = src[index[p]] for `width` floats per particle. */
int dsph_gather_f32(const float* src, const int32_t* index, int64_t n,
                    int32_t width, float* dst, void* stream);
/* Up to 8 per-particle arrays permuted by one index in one pass:
 * dst[f][p] = src[f][index[p]], words[f] 4-byte words per element (1, 2 or
 * 4; the arrays aligned to the element size).  src, dst and words are host
 * arrays of nfields entries.  No counterpart in the reference. */
int dsph_gather_set(const int32_t* index, int64_t n, int32_t nfields,
                    const void* const* src, void* const* dst,
                    const int32_t* words, void* stream);
/* dst[index[p]] = src[p] */
int dsph_scatter_f32(const float* src, const int32_t* index, int64_t n,
                     int32_t width, float* dst, void* stream);

/* Materialise the per-pair arrays of the reference's neighborDict
 * (v2/modules/neighborhood.py:433-460, 560-591) for a CSR list whose rows
 * are sorted query slots: i,j are ORIGINAL indices (int64), rows emitted in
 * original query order (row of original query q = qrank[q]); within a row
 * the order is the CSR order.  distances = r/h, vectors = d/(r+1e-7),
 * supports = h, kernels = W, gradients = grad W (Wendland C2, 2-D).
 * out_offsets[nq+1] are the CSR offsets in original order.  Any output
 * pointer may be NULL. */
int dsph_pairs_export(const dsph_pairs* pairs, const int32_t* q_sort_index,
                      const int32_t* q_rank, const int32_t* r_sort_index,
                      int32_t* out_offsets, int64_t* out_i, int64_t* out_j,
                      float* out_distances, float* out_vectors,
                      float* out_supports, float* out_kernels,
                      float* out_gradients, void* workspace,
                      size_t workspace_bytes, void* stream);
size_t dsph_pairs_export_workspace(int64_t nq);

/* ---------------------------------------------------------------------- */
/* SPH pair operators over a dsph_pairs list (kernel values on the fly)     */
/* ---------------------------------------------------------------------- */

/* rho_i = sum_j m_j W_ij (v2/sphOps.py:19-29, v2/modules/density.py:6-15).
 * Also writes vol_i = m_i/rho_i when vol != NULL (the reference's
 * 'actualArea', v2/modules/dfsph.py:59). `qmass` may be NULL when vol is. */
int dsph_density(const dsph_pairs* pairs, const float* rmass,
                 const float* qmass, float* rho, float* vol, void* stream);

/* dsph_density that also stores W_ij of every list entry in the list's own
 * layout (pair_w, as long as nbr), for later passes that need the kernel
 * values (the learned warm start's features).  No counterpart in the
 * reference. */
int dsph_density_w(const dsph_pairs* pairs, const float* rmass,
                   const float* qmass, float* rho, float* vol, float* pair_w,
                   void* stream);

/* Generic operator of v2/sphOps.py:253-282 over on-the-fly kernels:
 * op = 0 interpolate  sum_j m_j/rho_j q_j W_ij
 *      1 gradient     (mode: 0 naive, 1 difference, 2 summation, 3 symmetric)
 *      2 divergence   (mode as above; q is [n,2])
 *      3 divergenceConsistent (sphOps.py:108-149): the divergence with
 *        m_j / rho_i in place of m_j / rho_j (needs qrho)
 * `width` = components of q (1 or 2).  out is [nq, width] (interpolate),
 * [nq, width, 2] (gradient), [nq] (divergence). */
enum { DSPH_OP_INTERPOLATE = 0, DSPH_OP_GRADIENT = 1, DSPH_OP_DIVERGENCE = 2, DSPH_OP_DIVERGENCE_CONSISTENT = 3 };
enum { DSPH_GRAD_NAIVE = 0, DSPH_GRAD_DIFFERENCE = 1, DSPH_GRAD_SUMMATION = 2, DSPH_GRAD_SYMMETRIC = 3 };
int dsph_sph_operation(const dsph_pairs* pairs, int32_t op, int32_t grad_mode,
                       const float* qmass, const float* rmass,
                       const float* qrho, const float* rrho,
                       const float* qq, const float* rq, int32_t width,
                       float* out, void* stream);

/* ---------------------------------------------------------------------- */
/* DFSPH (v2/modules/dfsph.py, v2/simulationSchemes/dfsph.py:148-227)       */
/* ---------------------------------------------------------------------- */

typedef struct dsph_dfsph_params {
  float dt;
  float rho0;
  float omega;            /* relaxation factor */
  float error_threshold;  /* on mean(clamp(r/rho0, min=-thr)) */
  int32_t min_iters;
  int32_t max_iters;
  int32_t clamp_pressure; /* p = max(p, 0) after every update */
  int32_t source_density; /* 1: density source, 0: divergence source */
  int32_t check_every;    /* iterations launched between host checks (0 = 8) */
  int32_t scheme;         /* 0 DFSPH (v2/modules/dfsph.py); 1 ISPH relaxed Jacobi
                             (solveIncompressible_relaxedJacobi,
                             v2/modules/incompressible.py:308-364): pressure
                             gradient 'difference', error mean(|r|); the caller
                             passes alpha/dt, omega 0.5, no clamp, density source */
} dsph_dfsph_params;

typedef struct dsph_solve_stats {
  int32_t iterations;
  int32_t converged;       /* 1 when the loop ended below the threshold */
  double final_error;      /* reference error metric of the last iteration */
  double initial_error;    /* error metric of the first iteration */
} dsph_solve_stats;

/* Per-particle solver coefficients after summation density: for both
 * solves in one pair sweep (v2/modules/dfsph.py:7-33, 59-65):
 *   alpha_div = -dt   * V_i (|sum V_j gradW|^2 / m_i + sum V_j^2/m_j |gradW|^2)
 *   alpha_den = -dt^2 * (same bracket)
 *   div_i     = sum V_j (v_j - v_i) . gradW_ij    (divergence source)
 * and the delta-SPH inviscid viscosity acceleration
 * (v2/modules/viscosity.py:123-170) when visc != NULL.  areas are the rest
 * volumes, masses m = areas*rho0 in the alpha term as the reference. */
typedef struct dsph_visc_params {
  float alpha;   /* diffusion.alpha */
  float cs;      /* fluid.cs        */
  float rho0;
  float eps;     /* diffusion.eps   */
  int32_t pi_switch;
  int32_t scheme; /* 0 deltaSPH_inviscid, 1 deltaSPH_viscid (alpha = nu) */
} dsph_visc_params;

/* kgrad (optional, NULL = not stored): the scalar s_ij = V_j dW/dr /
 * (r + 1e-7) of every list entry, [entries] floats in the list's layout
 * (CSR or sliced).  The DFSPH coefficient V_j gradW_ij is d_ij * s_ij with
 * d_ij = x_i - x_j (min image), recomputed from the positions by every
 * sweep that streams kgrad; bit-identical to the on-the-fly value.  (ABI
 * version 3; version 2 stored the 8-byte vector.) */
int dsph_dfsph_prepare(const dsph_pairs* pairs, const float* vel,
                       const float* mass, const float* areas, const float* vol,
                       const float* rho, float dt, float rho0,
                       const dsph_visc_params* visc,
                       float* alpha_div, float* alpha_den, float* div_src,
                       float* visc_acc, float* kgrad, void* stream);

size_t dsph_dfsph_workspace(int64_t n);

/* Relaxed-Jacobi pressure solve (dfsphSolve, v2/modules/dfsph.py:52-109).
 * Inputs: vel, adv (advection acceleration; v* = vel + dt*adv), vol=m/rho,
 * areas, alpha (for this solve), src0 (for the divergence solve: the
 * divergence of vel; ignored for the density solve, whose source is
 * computed here as 1 - areas/vol - dt*div(v*)).  p: initial pressure on
 * entry (warm start), solution on exit.  accel: pressure acceleration of the
 * final p.  Convergence is decided on the device (fp64 mean residual);
 * the host is synchronised once every `check_every` iterations.  With kgrad
 * the loop keeps (x, y, p) and (x, y, a) records in the workspace
 * (dsph_pressure_accel_rec / dsph_pressure_update_rec). */
int dsph_dfsph_solve(const dsph_pairs* pairs, const dsph_dfsph_params* prm,
                     const float* kgrad, const float* vel, const float* adv, const float* rho,
                     const float* vol, const float* areas, const float* alpha,
                     const float* src0, float* p, float* accel,
                     float* residual, dsph_solve_stats* stats,
                     void* workspace, size_t workspace_bytes, void* stream);

/* dsph_dfsph_solve without a workspace argument and without any device
 * allocation.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The scratch is one static device array of
 * dsph_noalloc_scratch_bytes() (64 MiB) per GPU, held until the process
 * ends; it fits solves of up to dsph_dfsph_solve_noalloc_max_particles()
 * particles (1,676,400), larger ones return DSPH_ERR_CAPACITY before any
 * launch.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (two streams or host threads) race.  Results equal
 * dsph_dfsph_solve's bit for bit. */
size_t dsph_noalloc_scratch_bytes(void);
int64_t dsph_dfsph_solve_noalloc_max_particles(void);
int dsph_dfsph_solve_noalloc(const dsph_pairs* pairs,
                             const dsph_dfsph_params* prm, const float* kgrad,
                             const float* vel, const float* adv,
                             const float* rho, const float* vol,
                             const float* areas, const float* alpha,
                             const float* src0, float* p, float* accel,
                             float* residual, dsph_solve_stats* stats,
                             void* stream);

/* One pressure update (updatePressure, v2/modules/dfsph.py:39-50):
 * K_i = -dtf sum_j (a_j - a_i) . V_j gradW_ij with dtf = dt^2 for the
 * density solve (prm->source_density) else dt, p_i += omega (s_i - K_i) /
 * alpha_i (clamped at 0 when prm->clamp_pressure), residual_i = K_i - s_i
 * (may be NULL), and per 256-particle block the float64 sum of
 * max(residual_i / rho0, -error_threshold) into block_partials
 * [ceil(n/256)].  One pair sweep; dsph_dfsph_solve runs it in its loop. */
int dsph_pressure_update(const dsph_pairs* pairs, const dsph_dfsph_params* prm,
                         const float* kgrad, const float* vol, const float* src,
                         const float* alpha, const float* accel, float* p,
                         float* residual, double* block_partials, void* stream);

/* Pressure acceleration a_i = -(1/rho_i) sum_j V_j (p_i + p_j) gradW_ij
 * (computePressureAcceleration, v2/modules/dfsph.py:35-37); one pair sweep. */
int dsph_pressure_accel(const dsph_pairs* pairs, const float* kgrad,
                        const float* vol,
                        const float* rho, const float* p, float* accel,
                        void* stream);

/* The two sweeps of the solve loop in the form dsph_dfsph_solve runs them
 * when kgrad is stored: the gathered neighbor fields come as one 16-byte
 * record per particle, xp = (x, y, p, 0) and xa = (x, y, a_x, a_y) [n][4]
 * (x = the reference positions), so each pair costs one gather.
 * dsph_pack_pressure builds xp from p; dsph_pressure_accel_rec computes the
 * accelerations of dsph_pressure_accel into xa; dsph_pressure_update_rec
 * is dsph_pressure_update reading xa and writing p and xp.  Same values as
 * the unpacked entry points.  No counterpart in the reference (its arrays
 * are separate tensors). */
int dsph_pack_pressure(const dsph_pairs* pairs, const float* p, float* xp, void* stream);
int dsph_pressure_accel_rec(const dsph_pairs* pairs, const float* kgrad, const float* rho,
                            const float* xp, float* xa, void* stream);
int dsph_pressure_update_rec(const dsph_pairs* pairs, const dsph_dfsph_params* prm,
                             const float* kgrad, const float* src, const float* alpha,
                             const float* xa, float* p, float* xp, float* residual,
                             double* block_partials, void* stream);

/* Exactly `iterations` Jacobi iterations (sweep A + sweep B each, no
 * convergence test, no final acceleration) from the given p and source
 * (updatePressure, v2/modules/dfsph.py:39-50).  final_error (host, may be
 * NULL -> no synchronisation) receives the error metric of the last one.
 * accel is scratch: with kgrad the iterations keep the accelerations in the
 * workspace and accel is left unchanged. */
int dsph_dfsph_iterate(const dsph_pairs* pairs, const dsph_dfsph_params* prm,
                       const float* kgrad, const float* rho, const float* vol, const float* alpha,
                       const float* src, float* p, float* accel,
                       float* residual, int32_t iterations,
                       double* final_error, void* workspace,
                       size_t workspace_bytes, void* stream);

/* Error of a density field against rest density: out[0] = mean |rho/rho0-1|,
 * out[1] = max |rho/rho0-1| (fp64 reduction), copied to host. */
int dsph_density_error(const float* rho, int64_t n, float rho0, double* out2,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Semi-implicit Euler (v2/modules/integration.py:111-138):
 * v += dt*dudt; x += dt*v.  dudt = sum of up to 4 acceleration fields
 * (NULL entries skipped), as dfsph.py:209 a_p + g + visc + a_div. */
/* gravity2 is a HOST pointer to (gx, gy), or NULL. */
int dsph_integrate_semi_implicit(float* pos, float* vel, const float* a0,
                                 const float* a1, const float* a2,
                                 const float* a3, const float* gravity2,
                                 int64_t n, float dt, float* dudt_out,
                                 void* stream);

/* ---------------------------------------------------------------------- */
/* Generic SPH operation and DFSPH module pieces on a reference            */
/* neighborDict's per-pair arrays (v2/sphOps.py:253-303,                   */
/* v2/modules/dfsph.py:7-109)                                              */
/* ---------------------------------------------------------------------- */

/* A row-sorted pair list as the reference's neighborDict holds it:
 * pairs k of row r are offsets[r] .. offsets[r+1]-1; j[k] is the
 * neighbor's index in set B (int64, 'indices'[1]); per-pair arrays
 * 'kernels' W_ij [P], 'gradients' [P,2], 'distances' r_ij/h_ij [P],
 * 'vectors' x_ij [P,2], 'supports' h_ij [P], optional 'laplacians' [P]
 * (NULL where an operation does not read them). */
typedef struct dsph_pair_arrays {
  int64_t num_rows;         /* particles of set A                  */
  int64_t num_ref;          /* particles of set B                  */
  int64_t num_pairs;
  const int32_t* offsets;   /* [num_rows+1]                        */
  const int64_t* j;         /* [num_pairs]                         */
  const float* kernels;
  const float* gradients;
  const float* distances;
  const float* vectors;
  const float* supports;
  const float* laplacians;
} dsph_pair_arrays;

/* Row offsets of a pair list whose row indices `rows` (the reference's
 * indices[0], int64) are sorted ascending, as every reference neighbor
 * search returns them.  *unsorted (device int32) is cleared, then set when
 * a row index decreases or lies outside [0, num_rows); the offsets are then
 * not meaningful. */
int dsph_pp_offsets(const int64_t* rows, int64_t num_pairs, int32_t num_rows,
                    int32_t* offsets, int32_t* unsorted, void* stream);

/* sphOperation (v2/sphOps.py:253-282) on per-pair arrays.
 *   op:   0 density (sphDensityInterpolation, :19-29), 1 interpolate (:6-17),
 *         2 gradient (:32-62), 3 divergence (:65-106),
 *         4 divergenceConsistent (:108-149), 5 curl (:152-188),
 *         6 laplacian (:191-250)
 *   mode: gradientMode 0 naive, 1 difference, 2 summation, 3 symmetric;
 *         for the Laplacian also 4 conserving, 5 divergenceFree, 6 dot
 *         (anything else takes the Laplacian's default branch)
 *   dot_mode: divergenceMode 'dot' (contract the first axis of a 2x2
 *         tensor field) instead of 'div'.
 * The quantity is per particle (q_a [num_rows][q_width] for i, q_b
 * [num_ref][q_width] for j: the reference's (q_i, q_j) tuple; q_a may be
 * NULL where the mode reads q_j only) or per pair (q_pairs [P][q_width]:
 * the reference's single tensor).  q_width 1, 2 or 4 (a 2x2 tensor).
 * Output per row: density/curl 1, interpolate q_width, gradient
 * 2 q_width, divergence q_width/2, Laplacian q_width values. */
int dsph_pp_operation(const dsph_pair_arrays* pairs, int32_t op, int32_t mode,
                      int32_t dot_mode, const float* masses_a,
                      const float* masses_b, const float* densities_a,
                      const float* densities_b, const float* q_a,
                      const float* q_b, const float* q_pairs, int32_t q_width,
                      float* out, void* stream);

/* computeAlpha (v2/modules/dfsph.py:7-27): actual areas A = m / rho of
 * both sets, rest areas, dt_pow = dt^2 (density solve) or dt
 * (divergence solve). */
int dsph_pp_dfsph_alpha(const dsph_pair_arrays* pairs,
                        const float* actual_area_a, const float* actual_area_b,
                        const float* areas_a, const float* areas_b,
                        float dt_pow, float rho0, float* alpha, void* stream);

/* One updatePressure (v2/modules/dfsph.py:39-50) from the pressure
 * accelerations of both sets, with the convergence metric of :96 written to
 * *error_device (device float64: mean over rows of max(residual/rho0,
 * -threshold), summed in a fixed order). */
size_t dsph_pp_dfsph_update_workspace(int64_t num_rows);
int dsph_pp_dfsph_update(const dsph_pair_arrays* pairs, const float* masses_b,
                         const float* densities_b, const float* accel_a,
                         const float* accel_b, const float* source,
                         const float* alpha, const float* pressure_in,
                         float dt_pow, float omega, int32_t clamp_pressure,
                         float rho0, float threshold, float* pressure_out,
                         float* residual, double* error_device,
                         void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* Implicit particle shifting (computeShifting / solveShifting with scheme */
/* 'IPS', v2/modules/shifting.py:192-284, 318-460, 463-600)                */
/* ---------------------------------------------------------------------- */

/* The IPS system on a same-set neighbor list: rhs[n][2] (interleaved) =
 * sum_j A_j gradW_ij and, per list entry k (in the list's own layout, up
 * to hessian_entries), hessian[k] = A_j H_ij as (h00, h01, h11, 0) with the
 * Hessian2D of src/diffSPH/kernels.py:263-303 (Wendland C2).  The pair
 * geometry uses the reference's quotients. */
int dsph_ips_assemble(const dsph_pairs* pairs, const float* areas, float* rhs,
                      float* hessian, int64_t hessian_entries, void* stream);
/* y[i] = sum over row i of the entry blocks times x[j] (2-vectors per
 * particle, interleaved), the matvec of bicgstab_shifting. */
int dsph_ips_spmv(const dsph_pairs* pairs, const float* hessian,
                  int64_t hessian_entries, const float* x, float* y,
                  void* stream);
/* out = x + a y + b z over n floats (y, z may be NULL), float32 in that
 * order. */
int dsph_vec_lincomb(int64_t n, const float* x, float a, const float* y,
                     float b, const float* z, float* out, void* stream);
/* *out_device = sum a_k b_k (float64, fixed-order reduction); the largest
 * per-particle norm |(x_2i, x_2i+1)| of n particles. */
size_t dsph_vec_reduce_workspace(int64_t n);
int dsph_vec_dot(int64_t n, const float* a, const float* b, double* out_device,
                 void* workspace, size_t workspace_bytes, void* stream);
int dsph_vec_max_norm2(int64_t n, const float* x, double* out_device,
                       void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* Learned warm start of the DFSPH divergence solve (no counterpart in the  */
/* reference: its dfsphSolve warm-starts with 0.75 p_prev,                  */
/* v2/modules/dfsph.py:70-75)                                               */
/* ---------------------------------------------------------------------- */

/* Per-particle input features [n][16] (columns 8..15 zero):
 *   0 p_div_prev  1 sum_j V_j W_ij p_div_prev_j  2 s_div  3 sum V W s_div
 *   4 p_den_prev  5 sum_j V_j W_ij p_den_prev_j  6 rho/rho0-1  7 alpha_div
 * normalised as (f - mean16[c]) * istd16[c] (device arrays).  pair_w
 * (may be NULL): the W_ij of dsph_density_w for this list, streamed instead
 * of re-evaluating the pair geometry (same values).  records (may be NULL,
 * used with pair_w): scratch [n][4] floats; the neighbor fields are packed
 * there first so that an entry costs one gather (same values). */
int dsph_ws_features(const dsph_pairs* pairs, const float* vol,
                     const float* rho, const float* p_div_prev,
                     const float* p_den_prev, const float* div_src,
                     const float* alpha_div, float rho0, const float* pair_w,
                     const float* mean16, const float* istd16, float* X,
                     float* records, void* stream);

/* dsph_ws_features plus three second-ring features: columns 8, 9, 10
 * receive sum_j V_j W_ij t_j for the first ring's raw sums t = (column 1,
 * 3, 5 before normalisation), in the summation order of dsph_ws_features,
 * normalised like the others ((f - mean16[c]) * istd16[c] for c < nfeat,
 * 11 .. 16; 0 beyond).  pair_w is required; records and ring_records are
 * scratch [n][4] floats each (ring_records receives (V, raw sums) of the
 * first pass, the records of the second). */
int dsph_ws_features_ring(const dsph_pairs* pairs, const float* vol,
                          const float* rho, const float* p_div_prev,
                          const float* p_den_prev, const float* div_src,
                          const float* alpha_div, float rho0, const float* pair_w,
                          const float* mean16, const float* istd16, int32_t nfeat,
                          float* X, float* records, float* ring_records,
                          void* stream);

/* MLP 16 -> 64 -> 64 -> 64 -> 1 (ReLU), bf16 weights and layer inputs on
 * MFMA with fp32 accumulation.  Parameters are a flat fp32 array of
 * dsph_mlp_num_params() values: W1[16][64] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] w4[64] b4[1].  dsph_mlp_pack converts them to the
 * bf16 MFMA operand layout (dsph_mlp_packed_bytes() bytes) used by
 * forward and training.  dsph_mlp_forward: X [n][16] fp32, 16-byte aligned
 * (rows are read as float4; a misaligned X returns DSPH_ERR_ARG). */
int dsph_mlp_num_params(void);
int dsph_mlp_num_inputs(void);
size_t dsph_mlp_packed_bytes(void);
int dsph_mlp_pack(const float* params, void* packed, void* stream);
int dsph_mlp_forward(const float* X, int64_t n, const float* params,
                     const void* packed, float* Y, void* stream);
/* dsph_mlp_forward with its output mapped in the same pass:
 * Y = base + (y * scale + shift) (fp32, in that order), then negative
 * values set to 0 when clamp (NaN stays NaN); the warm-start pressure
 * p_prev + de-normalised network output.  No counterpart in the reference. */
int dsph_mlp_forward_affine(const float* X, int64_t n, const float* params,
                            const void* packed, const float* base, float scale,
                            float shift, int32_t clamp, float* Y, void* stream);
/* loss = sum_m w (y_m - target_m)^2 and its gradient w.r.t. every
 * parameter (grad: dsph_mlp_num_params() floats, deterministic fixed-order
 * reduction).  loss may be NULL. */
size_t dsph_mlp_train_workspace(int64_t n);
int dsph_mlp_loss_grad(const float* X, const float* target, int64_t n,
                       float loss_weight, const float* params,
                       const void* packed, float* grad, float* loss,
                       void* workspace, size_t workspace_bytes, void* stream);
/* Adam on fp32 master parameters (step >= 1 for the bias correction). */
int dsph_adam_step(float* params, const float* grad, float* m1, float* m2,
                   int64_t n, float lr, float beta1, float beta2, float eps,
                   float weight_decay, int32_t step, void* stream);

/* ---------------------------------------------------------------------- */
/* delta-SPH right-hand side, RK4 and time step                             */
/* ---------------------------------------------------------------------- */

/* Map neighbor entries from sorted slots to the query set's own indices:
 * nbr[e] = map[nbr[e]] for 0 <= nbr[e] < nmap, padding (-1) and
 * out-of-range entries become -1.  Used when the rows are queried in the
 * state's order against a grid sorted on the same positions (the
 * reference maps j through sortIndex the same way, neighborhood.cpp:253-435
 * and v2/modules/neighborhood.py:526). */
int dsph_remap_neighbors(int32_t* nbr, int64_t entries, const int32_t* map,
                         int64_t nmap, void* stream);

/* Parameters of one delta-SPH right-hand side (fluid only,
 * v2/simulationSchemes/deltaPlus.py:176-422). */
typedef struct dsph_deltasph_params {
  float rho0;
  float cs;            /* speed of sound */
  float delta;         /* density diffusion delta (densityDiffusion.py:73) */
  float kernel_scale;  /* h / (2 dx) (parameters.py:205) */
  float visc_coeff;    /* nu (viscid) or alpha (inviscid) */
  float visc_eps;      /* eps of pi_ij = v.x / (r + eps h^2) */
  float eos_kappa;     /* Tait kappa (eos == 1) */
  float gravity[2];    /* constant body acceleration (gravity.py:26-29) */
  int32_t viscosity;   /* -1 none, 0 deltaSPH_inviscid, 1 deltaSPH_viscid */
  int32_t pi_switch;   /* viscosity only for approaching pairs */
  int32_t eos;         /* 0 isoThermal p = cs^2 (rho - rho0), 1 Tait p = kappa (rho - rho0) */
  int32_t diffusion;   /* 0 off, 1 deltaSPH density diffusion */
  int32_t pad_;
} dsph_deltasph_params;

/* Wall particles seen by the fluid: `pairs` has the fluid particles as rows
 * (same nq and row order as the fluid list) and boundary indices as
 * entries (rpos = boundary positions); mass, rho (the mDBC densities of
 * dsph_mdbc), vel [nb,2]; the wall viscosity of computeViscosity_Boundary
 * (v2/modules/boundaryViscosity.py:123-218, config boundaryDiffusion). */
typedef struct dsph_boundary {
  const dsph_pairs* pairs;
  const float* mass;
  const float* rho;
  const float* vel;
  float visc_coeff;    /* nu (viscid) or alpha (inviscid) */
  float visc_eps;
  int32_t viscosity;   /* -1 none, 0 deltaSPH_inviscid, 1 deltaSPH_viscid */
  int32_t pi_switch;
} dsph_boundary;

/* Normalization matrices, their pseudo-inverses and the renormalized
 * density gradient of one particle set against itself:
 * M_i = sum_j V_j (x_j - x_i) (x) grad W_ij (identity for rows with < 4
 * entries), L_i = pinv(M_i) by the closed-form 2x2 SVD, and
 * grad^L rho_i = sum_j V_j (rho_i - rho_j) L_i grad W_ij, V_j = m_j/rho_j.
 * Replaces computeCovarianceMatrices (v2/modules/normalizationMatrices.py:
 * 18-23), pinv2x2 (v2/math.py:10-57), deltaPlus.py:245-248 and
 * renormalizedDensityGradient (v2/modules/densityDiffusion.py:7-26).
 * With `boundary` (may be NULL): M_i also sums the wall pairs
 * ('fluidwBoundary', deltaPlus.py:241) and the gradient adds the plain
 * wall term sum_b V_b (rho_i - rho_b) gradW_ib (densityGradient,
 * deltaPlus.py:262-263).
 * L [n,4] row-major and eigvals [n,2] may be NULL; grad_rho is [n,2]. */
int dsph_deltasph_normalization(const dsph_pairs* pairs, const dsph_boundary* boundary,
                                const float* mass, const float* rho, float* L,
                                float* eigvals, float* grad_rho, void* stream);

/* One fused pass: delta-SPH density diffusion (densityDiffusion.py:49-73),
 * viscosity (viscosity.py:123-170 / 200-218), continuity
 * (momentumEquation.py:6-39), EOS (pressureEOS.py:28-64) and pressure
 * acceleration (pressureForce.py:10-12), combined as deltaPlus.py:349-350:
 * dudt = a_p + g + visc [n,2], drhodt = continuity + diffusion [n].
 * With `boundary` (may be NULL) continuity and pressure add the wall pairs
 * (direction 'all', ftf + btf) and the viscosity adds the wall viscosity.
 * terms [n,8] (may be NULL) receives a_p.xy, visc.xy, continuity,
 * diffusion, pressure, 0 for inspection. */
int dsph_deltasph_rhs(const dsph_pairs* pairs, const dsph_boundary* boundary,
                      const dsph_deltasph_params* prm, const float* vel,
                      const float* mass, const float* rho, const float* grad_rho,
                      float* dudt, float* drhodt, float* terms, void* stream);

/* vol[i] = mass[i] / rho[i] [n]: the fluid volume V_j = m_j / rho_j of the
 * pair terms above (deltaPlus.py forms it per pair from the neighborDict),
 * computed once per particle. */
int dsph_volume(int64_t n, const float* mass, const float* rho, float* vol, void* stream);

/* dsph_deltasph_normalization reading the fluid V_j from `vol` (from
 * dsph_volume at the same mass and rho; NULL: formed per pair) in place of
 * m_j, instead of dividing m_j by rho_j for every pair: the same quotient,
 * so the results are bit-identical, and the pass is bound by VALU issue,
 * where the division is ~10 instructions (a further gathered field costs
 * the right-hand side pass more than its division, so it keeps it). */
int dsph_deltasph_normalization_v(const dsph_pairs* pairs, const dsph_boundary* boundary,
                                  const float* mass, const float* rho, const float* vol,
                                  float* L, float* eigvals, float* grad_rho, void* stream);

/* Ghost particles of the wall (buildBoundaryGhostParticles,
 * v2/modules/mDBC.py:9-17): x_g = x_b + (d + max(d, -h)) n, d the signed
 * distance (negative inside the wall), n the normal pointing into it. */
int dsph_boundary_ghosts(const float* boundary_pos, const float* distances,
                         const float* normals, int64_t nb, float support,
                         float* ghost_pos, void* stream);

/* mDBC wall densities: rows of `ghost_pairs` are the ghosts (qpos = ghost
 * positions, one per wall particle, in wall order), entries fluid indices
 * (rpos = fluid positions).  shepard [nb] may be NULL.
 * variant 0 -- mDBCDensity (v2/modules/mDBC.py:50-136): Shepard density for
 *   1..5 neighbors, rho0 for none, else the first-order MLS extrapolation
 *   c0 + (x_b - x_g) . c[1:] with c = pinv(A_g) b (3x3 pseudo-inverse,
 *   singular values <= 3 eps_f32 sigma_max dropped as torch.linalg.pinv
 *   does); no clamp.
 * variant 1 -- the C++ mDBC op (mDBCDensityKernel, v2/cppSrc/mDBC.h:129-244,
 *   host mDBC, v2/cppSrc/mDBC.cpp:3-147): the same system solved by
 *   Cramer's rule with det + 1e-7, Shepard density for <= 5 neighbors,
 *   the result clamped from below at rho0. */
int dsph_mdbc(const dsph_pairs* ghost_pairs, const float* fluid_mass,
              const float* fluid_rho, const float* boundary_pos, float rho0,
              float support, int32_t variant, float* boundary_rho, float* shepard,
              void* stream);

/* Stage state of the reference's RK4 (v2/modules/integration.py:23-40,
 * 231-276): x = x0 + dxdt c, u = u0 + dudt c, rho = rho0 + drhodt c, each
 * field skipped when its output is NULL. */
int dsph_rk_update(int64_t n, const float* x0, const float* dxdt, float* x,
                   const float* u0, const float* dudt, float* u,
                   const float* rho0, const float* drhodt, float* rho,
                   float c, void* stream);

/* Stage derivatives for the RK4 combination; a field whose state pointer is
 * NULL in dsph_rk4_combine may leave its entries NULL. */
typedef struct dsph_rk4_stages {
  const float* dxdt[4];    /* [n,2] each */
  const float* dudt[4];    /* [n,2] each */
  const float* drhodt[4];  /* [n]   each */
} dsph_rk4_stages;

/* state += (k1 + 2 k2 + 2 k3 + k4)/6 * dt per field
 * (v2/modules/integration.py:286-307); the averaged derivatives are also
 * written to *_avg when not NULL (the reference keeps them as dxdt, dudt,
 * drhodt, :489-494). */
int dsph_rk4_combine(int64_t n, const dsph_rk4_stages* k, float* x, float* u,
                     float* rho, float dt, float* dxdt_avg, float* dudt_avg,
                     float* drhodt_avg, void* stream);

/* dsph_rk_update / dsph_rk4_combine with the time step read from device
 * memory (dt[0], float32) when the kernels run: c = fl32(dt * 0.5) with
 * `half`, else dt, as the host forms them.  For steps whose time step is
 * computed on the device (dsph_timestep_dev) and that are captured into a
 * graph; results are bit-identical to the host-step forms. */
int dsph_rk_update_dev(int64_t n, const float* x0, const float* dxdt, float* x,
                       const float* u0, const float* dudt, float* u,
                       const float* rho0, const float* drhodt, float* rho,
                       const float* dt, int32_t half, void* stream);
int dsph_rk4_combine_dev(int64_t n, const dsph_rk4_stages* k, float* x, float* u,
                         float* rho, const float* dt, float* dxdt_avg,
                         float* dudt_avg, float* drhodt_avg, void* stream);

/* delta+ particle shifting (deltaPlusShifting, v2/modules/shifting.py:
 * 119-147), applied as solveShifting does with scheme 'deltaSPH', one
 * iteration and no free-surface treatment (shifting.py:880-1018) and added
 * to the positions as runner.py:40-42 does.  Per fluid particle i over its
 * fluid entries (and, with `boundary`, its wall entries: the merged state
 * of solveShifting):
 *   S_i = sum_j (1 + R (W_ij / W_0)^n) m_j / (rho_i + rho_j) gradW_ij,
 *   W_0 = W(fl32(dx / support), support),
 *   u_i = -((-CFL Ma) ((h_i / kernel_scale) 2)^2) S_i, each component
 *         clamped to +-fl32(threshold dx),
 *   Ma  = max |vel| / cs (or `mach` when >= 0, computeMach False: 0.1),
 *   pos_i += (pos_i - u_i) - pos_i.
 * Wall particles do not move.  shift_out [nq,2] (may be NULL) receives the
 * added displacement.  `pos` is the state's (unwrapped) position array in
 * row order; the pair geometry uses pairs->qpos / rpos.  (W/W_0)^n is the
 * float32 rounding of a double product chain for integral n in [0, 16]
 * (powf otherwise); the reference's torch.pow may differ by an ulp. */
typedef struct dsph_shift_params {
  double dx;           /* particle spacing (config particle.dx)            */
  double support;      /* config particle.support                          */
  float kernel_scale;
  float R;             /* shifting.R, default 0.25                         */
  float n;             /* shifting.n, default 4                            */
  float cfl;           /* shifting.CFL, default 1.5                        */
  float cs;
  float threshold;     /* shifting.threshold, default 0.5                  */
  float mach;          /* < 0: max |vel| / cs                              */
  int32_t pad_;
} dsph_shift_params;

size_t dsph_shift_workspace(int64_t nq);
int dsph_deltaplus_shift(const dsph_pairs* pairs, const dsph_boundary* boundary,
                         const dsph_shift_params* prm, const float* mass,
                         const float* rho, const float* vel, float* pos,
                         float* shift_out, void* workspace,
                         size_t workspace_bytes, void* stream);

/* max_i |v_i| of an [n,2] field (float32 norm).  Writes the host float
 * *out_host (synchronises the stream).  The workspace size serves both
 * reductions below. */
size_t dsph_max_norm_workspace(void);
int dsph_max_norm2(const float* v, int64_t n, float* out_host, void* workspace,
                   size_t workspace_bytes, void* stream);

/* sum of squares of n floats in float64, NaN entries skipped, in a fixed
 * reduction order.  The acceleration condition of computeTimestep
 * (v2/modules/timestep.py:22-25) takes torch.linalg.norm of
 * dudt[~isnan(dudt)], a flattened 1-D tensor: the norm of all components,
 * i.e. sqrt of this sum.  Writes *out_host (synchronises the stream). */
int dsph_sum_squares(const float* v, int64_t n, double* out_host, void* workspace,
                     size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* ISPH (v2/modules/incompressible.py, v2/sparse.py)                        */
/* ---------------------------------------------------------------------- */

/* Pressure Laplacian of projectVelocities (computeLaplacian,
 * v2/modules/incompressible.py:102-133), one value per list entry in the
 * list's own layout (values[k] belongs to nbr[k]): off-diagonal
 * t_ij = ((2 dt) (m_j / rho_j)) (|gradW_ij| / (q_ij h + 1e-7 h)), the
 * row's self entry -a_ii with a_ii = sum_j t_ij.  aii [n] and the Jacobi
 * preconditioner precond = -1 / a_ii (incompressible.py:178) may be NULL.
 * Replaces the COO tuple (A_, (i, j), n). */
int dsph_isph_laplacian(const dsph_pairs* pairs, const float* mass, const float* rho,
                        float dt, float* values, float* aii, float* precond, void* stream);

/* y = A x with A in the list layout (matvec_sparse_coo, v2/sparse.py:8-10). */
int dsph_spmv(const dsph_pairs* pairs, const float* values, const float* x, float* y,
              void* stream);

typedef struct dsph_bicgstab_stats {
  int32_t iterations;      /* completed iterations */
  int32_t status;          /* 0 |r| < atol, 1 |s| < atol (both converged), -10 rho
                              breakdown, -11 omega / (rtilde.v) breakdown, 2 maxiter */
  double residual_norm;    /* |r| (|s| after the s-test) at the last test (float64) */
  double rhs_norm;         /* |b| */
} dsph_bicgstab_stats;

/* Preconditioned BiCGStab of v2/sparse.py:141-230 (the scipy algorithm):
 * solves A x = b with M = diag(precond), x0 = x0_scale * x (in place),
 * stopping at |r| < max(atol, rtol |b|), the s-test, a breakdown or
 * maxiter.  fp64 != 0 runs the vectors in float64 (projectVelocities'
 * solverPrecision = float64, incompressible.py:173-200: A, b, M and x0
 * promoted, x rounded back); else float32 as the reference's default.
 * Dot products are float64 in both cases, with a fixed reduction order;
 * the loop control runs on the device and the host reads it every
 * check_every iterations (0 = 8).  Synchronises the stream. */
size_t dsph_bicgstab_workspace(int64_t n);
int dsph_bicgstab(const dsph_pairs* pairs, const float* values, const float* precond,
                  const float* b, float* x, float x0_scale, double rtol, double atol,
                  int32_t maxiter, int32_t check_every, int32_t fp64,
                  dsph_bicgstab_stats* stats, void* workspace, size_t workspace_bytes,
                  void* stream);

/* Source of the divergence-free system (computeDivergenceSourceTerm,
 * incompressible.py:148-165): B = rho0 * divergenceConsistent('difference')
 * of v* = vel + adv dt (adv may be NULL), minus its mean (float64 mean). */
size_t dsph_isph_source_workspace(int64_t n);
int dsph_isph_source(const dsph_pairs* pairs, const float* vel, const float* adv,
                     const float* mass, const float* rho, float dt, float rho0, float* B,
                     void* workspace, size_t workspace_bytes, void* stream);

/* Inputs of the relaxed-Jacobi solve (incompressible.py:269, 335-337):
 * alpha = alpha_den / dt (alpha_den: dsph_dfsph_prepare's density alpha),
 * p = 0.5 p_prev (p_prev NULL -> 0).  The solve itself is dsph_dfsph_solve
 * with scheme = 1. */
int dsph_isph_jacobi_init(int64_t n, const float* alpha_den, float dt, const float* p_prev,
                          float* alpha, float* p, void* stream);

/* Closing passes of the ISPH step (incompressible.py:398-460):
 *   a = -grad_summation(p_star) / rho0, v' = v* + a dt (v* = vel_in + adv dt),
 *   shift = -dt^2 grad_symmetric(p_ss) / rho0,
 *   vel = v' + grad_difference(v') . shift, pos = pos + v' dt + shift.
 * vprime, shift [n,2] receive v' and the shift; vel may alias vel_in. */
int dsph_isph_close(const dsph_pairs* pairs, const float* mass, const float* rho,
                    const float* vel_in, const float* adv, const float* p_star,
                    const float* p_ss, float dt, float rho0, float* vprime, float* shift,
                    float* pos, float* vel, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched ensemble of independent periodic delta-SPH domains               */
/* ---------------------------------------------------------------------- */

/* The reference runs several domains only as a Python loop over batch ids
 * (v2/neighborhoodFallback/neighborhood.py:121-183); each domain's step is
 * deltaPlus.simulationStep (v2/simulationSchemes/deltaPlus.py:176-422)
 * integrated with RK4 (v2/modules/integration.py:231-307) and
 * computeTimestep (v2/modules/timestep.py:8-39).  Here every pass of the
 * step runs over all domains in one launch: domain d owns the particles
 * [first_d, first_d + count_d) of the concatenated state, its own periodic
 * box, cell grid (global cell index = the domain's cell offset + local
 * cell) and time step.  Each domain's result is bit-identical to stepping
 * it alone with the single-domain entry points (same grid, same search
 * order, same float32 expressions and the same float64 reduction order of
 * the time-step norm). */

/* computeTimestep constants: dt_visc = 0.125 h^2 / nu / kernel_scale^2 and
 * dt_acoustic = cfl h / cs / kernel_scale, formed by the caller in float32
 * in that order; the acceleration condition
 * 0.25 sqrt(h / (|dudt| + 1e-7)) / kernel_scale (|dudt| = norm of all
 * components of the domain's RK4-averaged dudt) is evaluated per domain on
 * the device. */
typedef struct dsph_timestep_params {
  float support;
  float kernel_scale;
  float dt_visc;
  float dt_acoustic;
  float max_dt;
  float min_dt;
  int32_t adaptive;
  int32_t viscosity_constraint;
  int32_t acceleration_constraint;
  int32_t acoustic_constraint;
} dsph_timestep_params;

/* computeTimestep (v2/modules/timestep.py:8-39) of one domain on the
 * device, after an RK4 step: time[0] (float64, may be NULL) += dt[0], then,
 * when tp->adaptive, dt[0] = the new step from the RK4-averaged dudt_avg
 * [n,2].  accel_norm 0: |dudt| = sqrt of the float64 sum of squares of all
 * 2n components (the reduction tree of dsph_sum_squares); 1: the largest
 * per-particle norm (that of dsph_max_norm2).  The float32 formula and its
 * order are those of the host computation: no host synchronisation, the
 * result is bit-identical to reading the norm back and forming the step on
 * the host.  ws: dsph_max_norm_workspace() bytes. */
int dsph_timestep_dev(const float* dudt_avg, int64_t n, const dsph_timestep_params* tp,
                      int32_t accel_norm, float* dt, double* time, void* ws,
                      size_t ws_bytes, void* stream);

/* A batch laid out in one device workspace by dsph_batch_init. */
typedef struct dsph_batch {
  int64_t n;          /* particles over all domains */
  int64_t ncell;      /* cells over all domains */
  int32_t ndom;
  int32_t width;      /* neighbor entries per row (fixed-width sliced ELL) */
  int32_t block;      /* workgroup size of the per-particle kernels: 64, 128 or 256 */
  float support;      /* uniform h of every domain */
  float hsearch;      /* search radius the cell grids are sized for (>= support) */
  int32_t pad_;
  void* workspace;
  size_t workspace_bytes;
} dsph_batch;

/* Workspace for `ndom` periodic boxes with the given particle counts
 * (host arrays), grids sized for `hsearch`, `width` entries per row.
 * Returns 0 on invalid input. */
size_t dsph_batch_workspace(const dsph_domain* domains, const int64_t* counts, int32_t ndom,
                            float hsearch, int32_t width);

/* Lays the batch out in `workspace` and uploads the per-domain tables
 * (boxes, grids, offsets; synchronous).  All domains must be periodic on
 * both axes with at least 3 cells per axis; support <= hsearch. */
int dsph_batch_init(const dsph_domain* domains, const int64_t* counts, int32_t ndom,
                    float support, float hsearch, int32_t width, int32_t block,
                    void* workspace, size_t workspace_bytes, dsph_batch* out);

/* `steps` RK4 steps of every domain: sort into cell order, four right-hand
 * sides, RK4 combination, per-domain time step.  State arrays are the
 * concatenation of the domains (pos, vel [n,2], mass, rho [n], uid [n]),
 * dt [ndom] is the step each domain takes next (updated when adaptive),
 * time [ndom] (float64) accumulates.  No host synchronisation and no
 * allocation: the call may be captured into a graph.  Rows longer than
 * `width` are truncated and counted (dsph_batch_status). */
int dsph_batch_step(const dsph_batch* batch, const dsph_deltasph_params* prm,
                    const dsph_timestep_params* tp, float* pos, float* vel, float* mass,
                    float* rho, int32_t* uid, float* dt, double* time, int32_t steps,
                    void* stream);

/* The same `steps` steps captured once into an executable HIP graph
 * (*graph); replay with dsph_graph_launch on any stream.  The graph keeps
 * the pointers it was captured with. */
int dsph_batch_graph(const dsph_batch* batch, const dsph_deltasph_params* prm,
                     const dsph_timestep_params* tp, float* pos, float* vel, float* mass,
                     float* rho, int32_t* uid, float* dt, double* time, int32_t steps,
                     void** graph);
int dsph_graph_launch(void* graph, void* stream);
int dsph_graph_destroy(void* graph);

/* Sorts the current positions into cells and builds the neighbor rows
 * once; *max_count_host = the longest row, *pairs_host (may be NULL) = the
 * number of pairs (synchronises).  Used to size `width`. */
int dsph_batch_probe(const dsph_batch* batch, const float* pos, int32_t* max_count_host,
                     int64_t* pairs_host, void* stream);

/* Re-runs one pass of the last stage of the last step on the data it left
 * in the workspace (same inputs, same outputs): pass 0 = neighbor rows,
 * 1 = normalization + density gradient, 2 = right-hand side.  For timing a
 * single kernel; no counterpart in the reference.  Only valid after
 * dsph_batch_step (or a graph launch) has run at least one step. */
int dsph_batch_replay_pass(const dsph_batch* batch, const dsph_deltasph_params* prm,
                           int32_t pass, void* stream);

/* status[0] = rows truncated to `width` since the last reset, status[1] =
 * the longest row seen (synchronises).  reset != 0 clears both after
 * reading (with a fill kernel). */
int dsph_batch_status(const dsph_batch* batch, int32_t* status_host, int32_t reset, void* stream);

/* ---------------------------------------------------------------------- */
/* Adjoint of the delta-SPH right-hand side                                 */
/* ---------------------------------------------------------------------- */

/* Vector-Jacobian product of one fluid-only delta-SPH right-hand side
 * (dsph_deltasph_rhs without walls) with respect to positions, velocities
 * and densities: given the cotangents of dx/dt (= u, may be NULL), dudt
 * [n,2] and drhodt [n], writes dl/dx [n,2], dl/du [n,2], dl/drho [n].  It
 * differentiates every pair term, including the normalization matrix
 * through dL = -L dM L and the renormalized density gradient; the neighbor
 * set, the time step and the masses are constants.  The reference obtains
 * these gradients by torch autograd through its modules
 * (v2/simulationSchemes/deltaPlus.py:176-422, v2/modules/integration.py:
 * 231-307); it has no hand-written adjoint.  Requires one uniform support
 * in scatter mode (symmetric neighbor sets): the kernels gather both sides
 * of every pair, no atomics.  grad_rho is the forward's
 * dsph_deltasph_normalization output at the same state. */
size_t dsph_deltasph_vjp_workspace(int64_t n);
int dsph_deltasph_rhs_vjp(const dsph_pairs* pairs, const dsph_deltasph_params* prm,
                          const float* vel, const float* mass, const float* rho,
                          const float* grad_rho, const float* cot_dxdt,
                          const float* cot_dudt, const float* cot_drhodt,
                          float* grad_pos, float* grad_vel, float* grad_rho_out,
                          void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* Drop-in ops with the reference's per-pair inputs                         */
/* ---------------------------------------------------------------------- */

/* The three drop-ins below take the reference's per-pair arrays.  The CSR
 * given by (numNeighbors, neighborOffset) is clipped to [0, num_pairs) and
 * neighbor indices outside [0, num_ref) are skipped, so inconsistent input
 * cannot make a kernel read out of bounds (the reference trusts them). */

/* sphDensity (partiBench/cppOp.cpp:5-106): out_i = sum_k m_j[j_k] W_k over
 * CSR row i.  masses_j has num_ref entries, indices_j / kernels num_pairs. */
int dsph_op_sph_density(const float* masses_j, int64_t num_ref,
                        const int64_t* indices_j, const float* kernels,
                        int64_t num_pairs, int32_t num_particles,
                        const int32_t* num_neighbors,
                        const int32_t* neighbor_offset, float* out,
                        void* stream);

/* sphInterpolation (partiBench/cppOp.cpp:110-222,
 * sphOperations/cppOp.cpp:112-132): out_i = sum_k m_j/rho_j q_j W_k for a
 * quantity of `width` (1..4) components.  (The reference's CUDA branch
 * computes the density instead -- partiBench/cppOp.cpp:160-169; this follows
 * the documented formula of partiBench/cppOp.h:90-125.) */
int dsph_op_sph_interpolation(const float* masses_j, const float* densities_j,
                              const float* quantities_j, int32_t width,
                              int64_t num_ref, const int64_t* indices_j,
                              const float* kernels, int64_t num_pairs,
                              int32_t num_particles,
                              const int32_t* num_neighbors,
                              const int32_t* neighbor_offset, float* out,
                              void* stream);

/* scatterAdd (sphOperations/cppOp.cpp:364-423): out_i[c] = sum over the CSR
 * segment of row i of input[k, c] (input has num_pairs rows). */
int dsph_op_scatter_add(const float* input, int32_t width, int64_t num_pairs,
                        int32_t num_particles, const int32_t* num_neighbors,
                        const int32_t* neighbor_offset, float* out,
                        void* stream);

/* viscosityKernel (partiBench/viscosityKernel.cpp:5-21, per-particle
 * viscosityFn partiBench/viscosityKernel.h:98-145): for row i over its CSR
 * segment (h = supports[k], r = distances[k] h, xhat = vectors[k]):
 *   pi = ((v_i - v_j) . xhat) / ((r + eps) + h^2)  (0 where the dot >= 0
 *   and pi_switch is set), V = m_j / (rho_i + rho_j),
 *   out_i = (alpha h_i cs rho0 / rho_i) sum V pi gradients[k].
 * A-side arrays (velocities_i [.,2], densities_i, supports_i) have at least
 * num_particles rows; B-side arrays (velocities_j, masses_j, densities_j)
 * num_ref rows; out is [num_particles, 2]. */
int dsph_op_viscosity(const float* velocities_i, const float* velocities_j,
                      const float* masses_j, const float* densities_i,
                      const float* densities_j, const float* supports_i,
                      int64_t num_ref, const int64_t* indices_j,
                      const float* supports, const float* gradients,
                      const float* distances, const float* vectors,
                      int64_t num_pairs, int32_t num_particles,
                      const int32_t* num_neighbors,
                      const int32_t* neighbor_offset, float cs, float rho0,
                      float alpha, float eps, int32_t pi_switch, float* out,
                      void* stream);

/* mDBC (v2/cppSrc/mDBC.cpp:3-147, mDBCDensityKernel v2/cppSrc/mDBC.h:129-244)
 * on the caller's per-pair arrays: boundary particle b (row b of the CSR)
 * over its fluid neighbors j with x_bj = -(distances[k] supports[k])
 * vectors[k]; Shepard density sum m W / (sum V W + 1e-7); with more than 5
 * entries in the row, Cramer's rule on the 3x3 moment system (det + 1e-7)
 * extrapolated by (positions_boundary[b] - positions_ghost[b]); clamped below
 * at rho0.  masses_j / densities_j have num_ref entries; the position arrays
 * and both outputs num_boundary rows. */
int dsph_op_mdbc(const float* masses_j, const float* densities_j,
                 int64_t num_ref, const float* positions_boundary,
                 const float* positions_ghost, int32_t num_boundary,
                 const int64_t* indices_j, const float* kernels,
                 const float* gradients, const float* distances,
                 const float* vectors, const float* supports,
                 int64_t num_pairs, const int32_t* num_neighbors,
                 const int32_t* neighbor_offset, float rho0,
                 float* mdbc_density, float* shepard_density, void* stream);

/* computeNeighborhood, C++ form (v2/cppSrc/computeNeighborhood.cpp:5-94,
 * countingKernel / updateKernel v2/cppSrc/computeNeighborhood.h:115-187):
 * keeps the CSR entries of row i whose periodic offset d (per axis, in
 * double: ((p-q + L/2) - floor((p-q + L/2)/L) L) - L/2 with L = max - min in
 * float) has |d|^2 <= h_j^2.  min_domain, max_domain ([2] float) and
 * periodicity ([2] bool) are DEVICE pointers, as the reference reads them.
 * Count: counts[i] (i < num_particles) = kept entries of row i.
 * Fill: entry n of row i goes to new_offsets[i] + n (entries at or beyond
 * `capacity` are dropped) with out_i = i, out_j = j, rij = |d| / h_j,
 * xij = d / (|d| + 1e-6 h_i), hij = h_j. */
int dsph_op_neighborhood_count(const float* positions_i, const float* supports_i,
                               int32_t num_particles, const float* positions_j,
                               const float* supports_j, int64_t num_ref,
                               const int64_t* indices_j, int64_t num_pairs,
                               const int32_t* num_neighbors,
                               const int32_t* neighbor_offset,
                               const float* min_domain, const float* max_domain,
                               const uint8_t* periodicity, int32_t* counts,
                               void* stream);
int dsph_op_neighborhood_fill(const float* positions_i, const float* supports_i,
                              int32_t num_particles, const float* positions_j,
                              const float* supports_j, int64_t num_ref,
                              const int64_t* indices_j, int64_t num_pairs,
                              const int32_t* num_neighbors,
                              const int32_t* neighbor_offset,
                              const float* min_domain, const float* max_domain,
                              const uint8_t* periodicity,
                              const int32_t* new_offsets, int64_t capacity,
                              int64_t* out_i, int64_t* out_j, float* out_rij,
                              float* out_xij, float* out_hij, void* stream);

/* computeNeighborhood, Python form (v2/modules/neighborhood.py:433-460) with
 * the kernel values of neighborSearch (:585-589): the pair filter of the
 * Verlet path.  Positions are the WRAPPED positions (getPeriodicPosition);
 * the offset is the float32 min-image of v2/math.py:4-6 over `domain`,
 * h_ij by `mode` (DSPH_MODE_*), kept when |d| / h_ij <= 1; outputs rij =
 * |d| / h_ij, xij = d / (|d| + 1e-7), hij, and, where not NULL, the Wendland
 * C2 kernels [P] and gradients [P,2].  Rows, clipping and capacity as in
 * dsph_op_neighborhood_*. */
int dsph_pairs_filter_count(const dsph_domain* domain, int32_t mode,
                            const float* positions_i, const float* supports_i,
                            int32_t num_particles, const float* positions_j,
                            const float* supports_j, int64_t num_ref,
                            const int64_t* indices_j, int64_t num_pairs,
                            const int32_t* num_neighbors,
                            const int32_t* neighbor_offset, int32_t* counts,
                            void* stream);
int dsph_pairs_filter_fill(const dsph_domain* domain, int32_t mode,
                           const float* positions_i, const float* supports_i,
                           int32_t num_particles, const float* positions_j,
                           const float* supports_j, int64_t num_ref,
                           const int64_t* indices_j, int64_t num_pairs,
                           const int32_t* num_neighbors,
                           const int32_t* neighbor_offset,
                           const int32_t* new_offsets, int64_t capacity,
                           int64_t* out_i, int64_t* out_j, float* out_rij,
                           float* out_xij, float* out_hij, float* out_kernels,
                           float* out_gradients, void* stream);

/* Verlet reuse test inputs (v2/modules/neighborhood.py:462-478, 505-517):
 * out_host[0] = max_k |pos_now[k] - pos_prior[k]| over n particles,
 * out_host[1] = min of supports[0..num_supports).  Synchronises the stream
 * (the caller decides 2 out[0] > (verletScale - 1) out[1] on the host, as
 * the reference does).  Workspace: dsph_verlet_check_workspace(). */
size_t dsph_verlet_check_workspace(void);
int dsph_verlet_check(const float* pos_prior, const float* pos_now, int64_t n,
                      const float* supports, int64_t num_supports,
                      float* out_host, void* workspace, size_t workspace_bytes,
                      void* stream);

/* countNeighbors / buildNeighborList (neighborhood.cpp:65-227, 253-435;
 * device code neighborhood.h:100-325) over the caller's compact hash map:
 * hash_table [hash_map_length, 2] int32 (begin, length into cell_table or
 * -1), cell_table [num_cell_rows, 3] int64 (linear cell index, begin,
 * length into the sorted arrays), num_cells [2] int32, min/max domain [2]
 * float, periodicity [2] bool -- all DEVICE arrays.  For query i the cell
 * floor((x_i - min) / h_cell) and its (2 search_range + 1)^2 neighbors
 * (x offset outer, y inner; periodic axes wrapped) are looked up; sorted
 * particle j is a neighbor when |d| < h_j (scatter), < h_i (gather) or
 * < (h_i + h_j) / 2 (symmetric), |d| = sqrt of the float sum of the
 * double per-axis offsets.  The bucket of a cell is the non-negative
 * remainder of the int32 hash (73856093 cx + 19349663 cy) by
 * hash_map_length (the bucket the table builder files the cell under).
 * Count writes counts[num_query]; build writes neighbor i / j (int32, j in
 * sorted order) from offsets[i], dropping entries at or beyond capacity.
 * query_supports may be NULL in scatter mode, sorted_supports in gather
 * mode. */
int dsph_hash_count_neighbors(const float* query_positions,
                              const float* query_supports, int32_t num_query,
                              int32_t search_range, const float* sorted_positions,
                              const float* sorted_supports, int32_t num_sorted,
                              const int32_t* hash_table, int32_t hash_map_length,
                              const int32_t* num_cells, const int64_t* cell_table,
                              int64_t num_cell_rows, float h_cell,
                              const float* min_domain, const float* max_domain,
                              const uint8_t* periodicity, int32_t mode,
                              int32_t* counts, void* stream);
int dsph_hash_build_neighbors(const int32_t* offsets, int64_t capacity,
                              const float* query_positions,
                              const float* query_supports, int32_t num_query,
                              int32_t search_range, const float* sorted_positions,
                              const float* sorted_supports, int32_t num_sorted,
                              const int32_t* hash_table, int32_t hash_map_length,
                              const int32_t* num_cells, const int64_t* cell_table,
                              int64_t num_cell_rows, float h_cell,
                              const float* min_domain, const float* max_domain,
                              const uint8_t* periodicity, int32_t mode,
                              int32_t* out_i, int32_t* out_j, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DIFFSPH_AMD_H */
