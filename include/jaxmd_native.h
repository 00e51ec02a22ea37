/*
 * jaxmd_native.h - C ABI of the gfx950 neighbour-list MD library
 * (libjaxmd_native.so).
 *
 * Every entry point takes raw device pointers, sizes and plain attributes,
 * enqueues its kernels on the given HIP stream (``void*`` = hipStream_t,
 * NULL = default stream), performs no host synchronisation unless stated, and
 * returns an ``int`` status (JMD_OK on success).  No C++ exception crosses the
 * boundary.  Numerical overflow of a neighbour or cell list is a data output
 * (the uint8 partition error word, as in the reference), never a status.
 *
 * Buffers are owned by the caller.  Scratch memory comes in through an
 * explicit ``workspace`` pointer whose size is given by the matching
 * ``*_workspace_bytes`` query.
 *
 * Positions are [n, dim] float32, row-major (the reference's layout), with
 * dim 2 or 3, passed with a row stride ``position_stride`` in floats: dim
 * for a contiguous array, or 4 for 3-D rows padded to 16 bytes (base
 * 16-byte aligned), which the gather-heavy kernels read with one load.
 * A periodic space is described by ``periodic`` = 1 and ``side`` (host
 * pointer, dim float32 values: the float32 value of the side passed to
 * space.periodic); ``periodic`` = 0 is space.free(); ``periodic`` = 2 is
 * space.periodic_general with fractional coordinates (space.py:332-472):
 * positions are fractional in [0, 1), ``side`` holds the [dim, dim] box
 * matrix T row-major (invertible), displacements are T wrap(u_a - u_b),
 * forces and virials are in real space, the integrators' shift is
 * mod(u + T^-1 dR, 1), and cell grids are on the unit cube.  Entry points
 * that do not take periodic = 2 (the x64 and adjoint calls)
 * return JMD_ERR_UNSUPPORTED for it.
 *
 * Positions in a periodic space may lie anywhere (space.periodic(side,
 * wrapped=False) leaves them unwrapped, and the reference's displacement
 * folds any difference).  A caller that knows every position lies in
 * [0, side] per dimension ([0, 1] for periodic = 2) -- positions written by a
 * wrapping shift -- may OR JMD_POSITIONS_IN_BOX into ``periodic``: the force
 * kernels then skip the fold for particles far from every face and take a
 * single fold elsewhere.  Without the bit every difference is folded to the
 * minimum image whatever the positions.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to the jax-md 0.2.29 source tree).
 */
#ifndef JAXMD_NATIVE_H_
#define JAXMD_NATIVE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum {
  JMD_OK = 0,
  JMD_ERR_INVALID = 1,     /* bad argument (shape, null pointer, range) */
  JMD_ERR_LAUNCH = 2,      /* a HIP call failed */
  JMD_ERR_UNSUPPORTED = 3  /* valid request outside what the library does */
};

/* NeighborListFormat (jax_md/partition.py:641-661). */
enum { JMD_DENSE = 0, JMD_SPARSE = 1, JMD_ORDERED_SPARSE = 2 };

/* ORed into JMD_DENSE for jmd_neighbor_list_build: Dense rows kept by the
 * Sparse formats' pair test wrap(R_i - R_j) (partition.py:990-995) instead
 * of Dense's two-sided test, so that row i read in order (entries j < i for
 * OrderedSparse) is exactly the Sparse list's run of sender i.  No
 * counterpart in the reference: the block loops run Sparse-format lists as
 * such rows. */
enum { JMD_DENSE_SPARSE_TEST = 0x10 };

/* PartitionErrorCode bits (jax_md/partition.py:494-518). */
enum {
  JMD_NEIGHBOR_LIST_OVERFLOW = 1,
  JMD_CELL_LIST_OVERFLOW = 2,
  JMD_CELL_SIZE_TOO_SMALL = 4,
  JMD_MALFORMED_BOX = 8
};

/* Pair potentials (jax_md/energy.py:125-173 soft_sphere, :246-272
 * lennard_jones with :534-580 multiplicative_isotropic_cutoff). */
enum { JMD_LENNARD_JONES = 0, JMD_SOFT_SPHERE = 1 };

/* How a pair parameter is given (jax_md/smap.py:697-823):
 * scalar; per-particle vector combined as 0.5 (p_i + p_j) (smap.py:836);
 * species matrix p[s_i, s_j] (smap.py:787-797). */
enum { JMD_PARAM_SCALAR = 0, JMD_PARAM_PER_PARTICLE = 1, JMD_PARAM_SPECIES = 2 };

/* Flag ORed into a ``periodic`` argument: every position lies in the box
 * (see the note on positions above).  No counterpart in the reference. */
enum { JMD_POSITIONS_IN_BOX = 0x100 };

typedef struct jmd_param {
  int32_t mode;          /* JMD_PARAM_* */
  float value;           /* scalar mode */
  const float* values;   /* device: [n] (per-particle) or [s*s] (species) */
} jmd_param;

typedef struct jmd_pair_potential {
  int32_t kind;          /* JMD_LENNARD_JONES or JMD_SOFT_SPHERE */
  jmd_param sigma;
  jmd_param epsilon;
  jmd_param alpha;       /* soft sphere exponent */
  float r_onset;         /* LJ: absolute onset radius (2.0 * max sigma) */
  float r_cutoff;        /* LJ: absolute cutoff radius (2.5 * max sigma);
                            soft sphere: an upper bound of sigma over all
                            pairs, or 0 when unknown (a hint only) */
  int32_t n_species;     /* species matrices are n_species x n_species */
  const int32_t* species;/* device [n] or NULL */
} jmd_pair_potential;

/* Library version string. */
const char* jmd_version(void);

/* ------------------------------------------------------------------------ */
/* Cell list and neighbour list                                              */
/* ------------------------------------------------------------------------ */

/* Scratch bytes needed by jmd_cell_list / jmd_neighbor_list_build for n
 * particles on a grid of cell_count cells (cell_count 0: no cell list) with
 * cell capacity cell_capacity (the largest capacity later builds pass; 0 for
 * calls without a list build).  The workspace must be zero-filled before its
 * first use; every call leaves it in that state again, except int32 word 4,
 * which counts the completed list builds (gated builds that ran, and
 * ungated ones). */
size_t jmd_neighbor_workspace_bytes(int64_t n, int32_t dim, int64_t cell_count,
                                    int32_t cell_capacity, int32_t format);

/* partition.cell_list(...).allocate/update body: cell_list_fn
 * (jax_md/partition.py:349-471).  Bins n particles into the
 * cells_per_side grid (host arrays of dim values, x fastest in the hash) with
 * per-cell capacity cell_capacity, writing the reference's buffers:
 *   position_buffer [cell_count*cap, dim] (zeros in empty slots),
 *   id_buffer       [cell_count*cap]      (n in empty slots),
 *   particle_cell_id[n, dim].
 * Slot of the k-th particle in (cell, id) order is cell*cap + (k mod cap)
 * (partition.py:441-442).  *filling_max (device int32) receives the largest
 * cell occupancy; *overflow (device uint8) is set to 1 when it exceeds
 * cell_capacity (partition.py:458-460). */
int jmd_cell_list(const float* position, int64_t n, int32_t dim,
                  int32_t position_stride, const int32_t* cells_per_side, const float* cell_size,
                  int32_t cell_capacity, float* position_buffer,
                  int32_t* id_buffer, int32_t* particle_cell_id,
                  int32_t* filling_max, uint8_t* overflow, void* workspace,
                  void* stream);

/* Largest cell occupancy on a grid, the quantity count_cell_filling +
 * _estimate_cell_capacity read on the host (jax_md/partition.py:191-208,
 * 243-250).  *filling_max is a device int32. */
int jmd_cell_filling_max(const float* position, int64_t n, int32_t dim,
                         int32_t position_stride, const int32_t* cells_per_side,
                         const float* cell_size,
                         int32_t* filling_max, void* workspace, void* stream);

/* Body of neighbor_fn in partition.neighbor_list
 * (jax_md/partition.py:1037-1117): candidates from the 3^dim stencil of the
 * cell grid (use_cell_list = 1, partition.py:911-951) or from all particles
 * (partition.py:904-909), self-masking (:953-958), and the prune of the
 * chosen format (:960-1032), written in the reference's order:
 *   Dense  idx [n, max_occupancy] int32, rows left-compacted, padding n;
 *   Sparse / OrderedSparse idx [2, max_occupancy] = (receiver, sender).
 * cutoff_sq is float32 (r_cutoff + dr_threshold)^2.  cell_capacity is the
 * reference cell capacity (it fixes the slot order of candidates and the
 * cell-overflow bit).  occupancy (device int32 [1]) receives the largest row
 * count (Dense) or the total pair count (Sparse); the neighbour-overflow bit
 * is OR-ed into *error when it exceeds max_occupancy, the cell-overflow bit
 * when a cell holds more than cell_capacity particles (:1066,1110).
 * reference_position [n, dim] (contiguous) receives a copy of position.
 * With error = NULL the call only measures (idx is not written and may be
 * NULL): occupancy receives the largest Dense row count (or the Sparse
 * total); this is how allocate sizes the list (partition.py:1092-1105).
 * gate (device int32, may be NULL): when non-NULL the whole call is a no-op
 * on the device unless *gate != 0, and *gate is reset to 0 at its end; this
 * is the lax.cond of update (:1146-1154) without a host round trip. */
int jmd_neighbor_list_build(const float* position, int64_t n, int32_t dim,
                            int32_t position_stride, int32_t periodic,
                            const float* side, float cutoff_sq,
                            int32_t use_cell_list, const int32_t* cells_per_side,
                            const float* cell_size, int32_t cell_capacity,
                            int32_t format, int32_t mask_self, int32_t* idx,
                            int64_t max_occupancy, int32_t* occupancy,
                            uint8_t* error, float* reference_position,
                            int32_t* gate, void* workspace, void* stream);

/* Skin criterion of neighbor_list update (jax_md/partition.py:1141-1154):
 * sets *gate (device int32) to 1 when any particle satisfies
 * metric_sq(position, reference_position) > threshold_sq, with threshold_sq
 * = float32 (dr_threshold / 2)^2.  *gate is only ever raised here; it is
 * lowered by the gated jmd_neighbor_list_build. */
int jmd_neighbor_list_check(const float* position, int32_t position_stride,
                            const float* reference_position, int64_t n,
                            int32_t dim, int32_t periodic, const float* side,
                            float threshold_sq, int32_t* gate, void* stream);

/* Cell order of the particles: order[s] is the id of the s-th particle when
 * sorted by cell hash (cells_per_side grid, x fastest) and by id inside a
 * cell, the order of the reference's cell list (partition.py:432);
 * inverse[order[s]] = s.  Either output may be NULL.  No counterpart call in
 * the reference: it is the spatial layout the native block loop runs in. */
int jmd_cell_order(const float* position, int64_t n, int32_t dim,
                   int32_t position_stride, const int32_t* cells_per_side,
                   const float* cell_size, int32_t* order, int32_t* inverse,
                   void* workspace, void* stream);

/* Row permutation: out[r] = in[perm[r]] (inverse = 0) or out[perm[r]] =
 * in[r] (inverse = 1) for n rows of row_bytes bytes with the given row
 * strides (all multiples of 4 bytes); in and out must not alias. */
int jmd_permute_rows(const void* in, void* out, const int32_t* perm, int64_t n,
                     int32_t row_bytes, int32_t in_stride_bytes,
                     int32_t out_stride_bytes, int32_t inverse, void* stream);

/* Relabels a neighbour list under a particle permutation: Dense
 * out[perm[r], k] = perm[in[r, k]]; Sparse / OrderedSparse entries keep their
 * place and each id j becomes perm[j] (OrderedSparse pairs are stored with
 * receiver < sender, swapped where the relabelling reversed them,
 * partition.py:998).  Padding n stays n. */
int jmd_remap_neighbor_ids(const int32_t* idx_in, int32_t* idx_out,
                           const int32_t* perm, int64_t n, int32_t format,
                           int64_t width, void* stream);

/* ------------------------------------------------------------------------ */
/* Pair reduction                                                            */
/* ------------------------------------------------------------------------ */

/* Scratch bytes for jmd_pair_energy_force. */
size_t jmd_pair_workspace_bytes(int64_t n);

/* Threads per Dense row in jmd_pair_energy_force (2, 4, 8 or 16; 0 for
 * the built-in choice).  A process-wide tuning knob set by the autotuner; it
 * has no counterpart in the reference (XLA picks its own fusion layout). */
int jmd_set_threads_per_atom(int32_t threads_per_atom);

/* smap.pair_neighbor_list(fn, metric, ...) applied to
 * energy.lennard_jones / energy.soft_sphere
 * (jax_md/smap.py:856-979, jax_md/energy.py:200-343), plus its gradient.
 * For idx in format `format` (idx_width = max_occupancy) it computes, each
 * output optional (NULL skips it):
 *   force          [n, dim] float32  = -dE/dR
 *   energy_atom    [n]      float32  per-particle energy (reduce_axis=(1,),
 *                                    Dense and Sparse only)
 *   energy_total   [1]      float64  total energy, summed in float64
 * Pair arithmetic is float32 with float64 per-particle accumulation.
 * When momentum is non-NULL (force required) the call also applies
 * p += half_dt * F (the closing half kick of simulate.velocity_verlet,
 * simulate.py:240-241): in the same pass for Dense lists, in a second
 * launch for the Sparse formats; dt_device (device float32 [1], may be
 * NULL) overrides half_dt with float32(dt / 2). */
int jmd_pair_energy_force(const float* position, int64_t n, int32_t dim,
                          int32_t position_stride, int32_t periodic,
                          const float* side,
                          const jmd_pair_potential* potential,
                          const int32_t* idx, int64_t idx_width,
                          int32_t format, float* force, float* energy_atom,
                          double* energy_total, float* momentum,
                          float half_dt, const float* dt_device,
                          void* workspace, void* stream);

/* jmd_pair_energy_force without a workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Arguments and results are those of jmd_pair_energy_force; the
 * float64 energy partials live in a fixed 1 MiB device array of the library
 * (one per GPU, held until the process ends), so n may not exceed
 * JMD_PAIR_NOALLOC_MAX_N: a larger n returns JMD_ERR_INVALID before anything
 * is enqueued.  The call allocates nothing (no hipMalloc, no library
 * allocation).  All calls on one GPU share that array: calls that overlap
 * on one GPU (different streams or threads without ordering) race. */
#define JMD_PAIR_NOALLOC_MAX_N 2000000
int jmd_pair_energy_force_noalloc(const float* position, int64_t n, int32_t dim,
                                  int32_t position_stride, int32_t periodic,
                                  const float* side,
                                  const jmd_pair_potential* potential,
                                  const int32_t* idx, int64_t idx_width,
                                  int32_t format, float* force, float* energy_atom,
                                  double* energy_total, float* momentum,
                                  float half_dt, const float* dt_device,
                                  void* stream);

/* The virial W_ab = dU/d eps_ab of the pair energy with every displacement
 * perturbed as (I + eps) d, i.e. weight * sum over the list's ordered pairs
 * of (U'(r)/r) d_a d_b (weight 1/2 for Dense and Sparse, 1 for
 * OrderedSparse): what quantity.stress and quantity.pressure obtain with
 * jax.grad through the `perturbation` argument of the displacement
 * (quantity.py:202-282, space.py:299-300).  Pair terms float32, sums
 * float64; virial: device float64 [dim, dim] (symmetric).  workspace:
 * jmd_pair_virial_workspace_bytes() bytes. */
size_t jmd_pair_virial_workspace_bytes(void);
int jmd_pair_virial(const float* position, int64_t n, int32_t dim, int32_t position_stride,
                    int32_t periodic, const float* side, const jmd_pair_potential* potential,
                    const int32_t* idx, int64_t idx_width, int32_t format, double* virial,
                    void* workspace, void* stream);

/* Per-entry pair energies: smap.pair_neighbor_list with reduce_axis=()
 * (smap.py:940-953), the energy of each list entry divided by the
 * reference's normalization (2 for Dense and Sparse, 1 for OrderedSparse),
 * 0 for padding entries.  out: device float32, [n, idx_width] for Dense
 * (row-major, like idx), [idx_width] for the Sparse formats.  Same
 * position, geometry, potential and list arguments as jmd_pair_energy_force;
 * no workspace. */
int jmd_pair_edge_energies(const float* position, int64_t n, int32_t dim,
                           int32_t position_stride, int32_t periodic, const float* side,
                           const jmd_pair_potential* potential, const int32_t* idx,
                           int64_t idx_width, int32_t format, float* out, void* stream);

/* The x64 form (the reference's JAX_ENABLE_X64 run): float64 positions,
 * forces, momenta and pair arithmetic, parameters float32-valued as
 * maybe_downcast leaves them (util.py:109-112).  Dense lists only
 * (format = JMD_DENSE); same potential, list and outputs as
 * jmd_pair_energy_force otherwise (smap.py:856-979, energy.py:246-272,
 * 534-580).  side: host float64 [dim].  energy_total needs a workspace of
 * jmd_pair_workspace_bytes_f64(n) bytes. */
size_t jmd_pair_workspace_bytes_f64(int64_t n);
int jmd_pair_energy_force_f64(const double* position, int64_t n, int32_t dim,
                              int32_t periodic, const double* side,
                              const jmd_pair_potential* potential,
                              const int32_t* idx, int64_t idx_width, int32_t format,
                              double* force, double* energy_total, double* momentum,
                              double half_dt, void* workspace, void* stream);

/* ------------------------------------------------------------------------ */
/* Adjoint of the pair force                                                 */
/* ------------------------------------------------------------------------ */

/* Scratch bytes for jmd_pair_force_vjp. */
size_t jmd_adjoint_workspace_bytes(int64_t n, int32_t dim);

/* Number of float64 values in the parameter outputs of jmd_pair_force_vjp
 * and jmd_adjoint_param_gather: [sigma block][epsilon block], each 1 value
 * for a scalar parameter or n for a per-particle one (diameters: p_ij =
 * (p_i + p_j) / 2, smap.py:836); -1 for species matrices (not
 * differentiated). */
int64_t jmd_adjoint_param_count(const jmd_pair_potential* potential, int64_t n);

/* Vector-Jacobian product of the force F(R) = -dE/dR of
 * jmd_pair_energy_force (Dense lists), the gradient the reference obtains
 * from jax.grad through smap.pair_neighbor_list (quantity.py:58-60,
 * smap.py:856-979).  For a cotangent v [n, dim]:
 *   position_cotangent [n, dim] = (dF/dR)^T v  (dF/dR is symmetric)
 *   param_cotangent    = v^T dF/dsigma, v^T dF/depsilon
 *   energy_param_grad  = dE/dsigma, dE/depsilon
 * the last two float64 in the layout of jmd_adjoint_param_count (scalar
 * parameters: [2]; a per-particle sigma gives n values dF/dsigma_i first);
 * NULL skips them; r_onset / r_cutoff are constants (energy.py:319-320).
 * Pair terms float32, row sums float64, fixed-order reductions; the
 * per-particle values are row sums of their own (no atomics). */
int jmd_pair_force_vjp(const float* position, int64_t n, int32_t dim,
                       int32_t position_stride, int32_t periodic, const float* side,
                       const jmd_pair_potential* potential, const int32_t* idx,
                       int64_t idx_width, int32_t format, const float* cotangent,
                       float* position_cotangent, double* param_cotangent,
                       double* energy_param_grad, void* workspace, void* stream);

/* Doubles in the parameter accumulator of jmd_nve_adjoint_step
 * (zero-filled by the caller): per-workgroup sums for scalar parameters,
 * per-particle sums for per-particle ones. */
size_t jmd_adjoint_param_acc_doubles(int64_t n, int32_t dim);

/* One backward step of the velocity-Verlet adjoint (simulate.py:227-243
 * differentiated as jax.grad would through lax.fori_loop): with h = J(x) ph_in,
 *   x_adjoint += coef_x h        (float64, [n, dim])
 *   ph_out     = ph_in + coef_p x_adjoint   (float32; must not alias ph_in)
 *   param_acc[2 b + k] += coef_param (ph_in^T dF/dtheta_k), k = sigma, eps
 * param_acc may be NULL.  No counterpart entry point in the reference. */
int jmd_nve_adjoint_step(const float* position, int64_t n, int32_t dim,
                         int32_t position_stride, int32_t periodic, const float* side,
                         const jmd_pair_potential* potential, const int32_t* idx,
                         int64_t idx_width, const float* ph_in, float* ph_out,
                         double* x_adjoint, float coef_x, float coef_p, double* param_acc,
                         float coef_param, void* stream);

/* jmd_nve_adjoint_step for an x64 trajectory: positions float64 [n, dim]
 * (each displacement formed in float64, rounded once; the product itself in
 * float32 as in jmd_nve_adjoint_step). */
int jmd_nve_adjoint_step_x64(const double* position, int64_t n, int32_t dim,
                             int32_t periodic, const float* side,
                             const jmd_pair_potential* potential, const int32_t* idx,
                             int64_t idx_width, const float* ph_in, float* ph_out,
                             double* x_adjoint, float coef_x, float coef_p,
                             double* param_acc, float coef_param, void* stream);

/* out[0..1] = fixed-order sums of the accumulator's two columns (scalar
 * sigma and epsilon). */
int jmd_adjoint_param_sum(const double* param_acc, int64_t n, int32_t dim, double* out,
                          void* stream);

/* The accumulated parameter cotangents in the layout of
 * jmd_adjoint_param_count (scalar parameters summed in a fixed order,
 * per-particle ones copied); param_acc keeps its values. */
int jmd_adjoint_param_gather(double* param_acc, int64_t n, int32_t dim,
                             const jmd_pair_potential* potential, double* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Ragged ensembles of 2-D systems (one persistent launch)                   */
/* ------------------------------------------------------------------------ */

/* One system of an ensemble: particles [offset, offset + n) of the
 * concatenated arrays, its periodic box, the cell grid of its neighbour list
 * (cells_per_side = floor(side / cutoff), cell_size = side / cells_per_side,
 * partition.py:146-188), the grid its cell capacity is counted on
 * (partition.py:249-250 derives it a second time), its capacities and list
 * radii, and where its species parameters start (a potential's species
 * matrices may hold one block per system). */
typedef struct {
  int32_t offset, n;
  int32_t cells_per_side[2];
  float cell_size[2];
  int32_t fill_cells_per_side[2];
  float fill_cell_size[2];
  float side[2];
  int32_t cell_capacity, max_occupancy;
  float cutoff_sq;      /* float32 (r_cutoff + dr_threshold)^2 of its list */
  float threshold_sq;   /* float32 (dr_threshold / 2)^2 */
  int32_t param_offset; /* offset of its species-parameter block (0: shared) */
  int32_t reserved;
} jmd_system2d;

/* Largest cell count (either grid) a system may have. */
int jmd_fire_ensemble_max_cells(void);

/* Threads per system workgroup of jmd_fire_ensemble_2d and its adjoint: 256,
 * 512 or 1024; 0 (the default) chooses 1024 when the systems average 4096
 * particles or more, else 512.  A process-wide tuning knob with no
 * counterpart in the reference; results differ only in the summation order
 * of the workgroup reductions. */
int jmd_fire_ensemble_set_threads(int32_t threads);

/* Soft-sphere FIRE minimisation of every system of a ragged ensemble, one
 * workgroup per system (systems[order[b]] for workgroup b): the reference's
 * loop `nbrs = nbrs.update(R); state = apply_fn(state, neighbor=nbrs)` of
 * minimize.fire_descent (minimize.py:124-226) over
 * energy.soft_sphere_neighbor_list with a Dense list (energy.py:200-243,
 * partition.py:801-1164), for every system at once.  mode:
 *   0  out_stat[s] = largest cell filling on the fill grid (for capacities)
 *   1  out_stat[s] = largest row count of a list built now (for widths)
 *   2  allocate: build the list (idx [n_total, width], global ids, padding
 *      n_total), F = F(R), P = 0, dt = dt_start, alpha = alpha_start,
 *      energy[s] = E(R)
 *   3  n_steps FIRE steps; energy[s] = E at the end
 * fire_params = static_cast float16 values (dt_start, dt_max, n_min, f_inc,
 * f_dec, alpha_start, f_alpha); fire_state [n_systems, 2] = (dt, alpha),
 * fire_n_pos [n_systems]; error [n_systems] sticky overflow bits; builds
 * [n_systems] counts list builds; scratch holds 3 * n_total int32.  The
 * potential's parameters index particles by their global (concatenated) id.
 * width is the row stride of idx (at least every system's max_occupancy); a
 * multiple of 4 with a 16-byte aligned idx lets the pair loops read rows 16
 * bytes at a time.
 * No counterpart entry point in the reference, which runs one system at a
 * time. */
int jmd_fire_ensemble_2d(const jmd_system2d* systems, const int32_t* order, int32_t n_systems,
                         int64_t n_total, int32_t mode, int32_t n_steps, double* position_hi,
                         float* position, float* momentum, float* force,
                         float* reference_position, int32_t* idx, int32_t width,
                         int32_t* scratch, const jmd_pair_potential* potential,
                         const float* fire_params, float* fire_state, int32_t* fire_n_pos,
                         uint8_t* error, int32_t* out_stat, double* energy, int32_t* builds,
                         float* tape_x, float* tape_force, float* tape_q, int32_t* tape_idx,
                         float* tape_scalars, void* stream);

/* Adjoint of jmd_fire_ensemble_2d's trajectory (no counterpart entry point in
 * the reference, which would take jax.grad through each system's
 * lax.fori_loop over fire_descent's apply_fn, minimize.py:124-226).  The
 * loss is scale * sum_s E_s(x_T); theta[s] is its derivative with respect to
 * system s's parameter theta_s, through which the species matrix sigma
 * depends with coefficients dsigma (d sigma_ab / d theta_s, laid out like
 * sigma).  FIRE's float16 dt / alpha / reset decisions are piecewise
 * constant and carry no derivative.  Mode 3 of jmd_fire_ensemble_2d with
 * tape_* set records a segment (per node k = 0..n_steps: float32 positions,
 * forces [n+1][n_total][2]; list slots [n+1][n_total][width], slot 0 the
 * first node's list and slot k + 1 a system's list when it rebuilt at step
 * k; per step: momenta before the mixing [n][n_total][2] and six scalars
 * per system [n][n_systems][6], the sixth the slot of the list of the
 * step's second node).  mode: 0 initialises x_adjoint = -scale F,
 * p_adjoint = pending = 0, theta = scale dE/dtheta at (position, idx);
 * 1 runs n_steps backward steps through a tape; 2 applies the pending
 * vector at the tape's first node (call once, after the earliest segment).
 * box_adjoint (optional, [n_systems] float64): the loss's derivative with
 * respect to each system's box side L, the particles held in real
 * coordinates (the meta-optimisation notebook's periodic_general(L(d) I,
 * fractional_coordinates=False) under jax.grad): L moves the minimum images
 * of the pair terms and the folds of the shift; mode 0 sets it, modes 1 and
 * 2 add to it.  NULL: the boxes are constants. */
int jmd_fire_ensemble_adjoint_2d(const jmd_system2d* systems, int32_t n_systems,
                                 int64_t n_total, int32_t mode, int32_t n_steps,
                                 const float* position, const float* force, const int32_t* idx,
                                 int32_t width, const jmd_pair_potential* potential,
                                 const float* dsigma, const float* tape_x,
                                 const float* tape_force, const float* tape_q,
                                 const int32_t* tape_idx, const float* tape_scalars,
                                 double* x_adjoint, double* p_adjoint, float* pending,
                                 float* scratch, double* theta, double* box_adjoint,
                                 float scale, void* stream);

/* ------------------------------------------------------------------------ */
/* Integrators                                                               */
/* ------------------------------------------------------------------------ */

/* First half of simulate.velocity_verlet (jax_md/simulate.py:227-239):
 *   p' = p + half_dt * F;  R' = shift(R, dt * p' / m)
 * with the position accumulated in float64 (position_hi [n, dim]) and
 * mirrored to the float32 position_out [n, dim] (row stride
 * position_stride).  Outputs may alias the
 * matching inputs (in-place step).  mass: NULL for the scalar mass_scalar,
 * else device [n].  The shift wraps into the box with jnp.mod semantics when
 * periodic (space.py:250-252).  When reference_position and gate are
 * non-NULL the skin criterion of jmd_neighbor_list_check is evaluated on the
 * new positions in the same pass (gate raised when any particle moved more
 * than sqrt(threshold_sq) since the last build).  dt_device (device float32
 * [1], may be NULL) overrides dt, with half_dt = float32(dt / 2) as in
 * velocity_verlet (simulate.py:235-236); FIRE keeps its dt on the device. */
int jmd_nve_kick_drift(const double* position_hi, const float* momentum,
                       const float* force, const float* mass,
                       float mass_scalar, int64_t n, int32_t dim,
                       int32_t periodic, const float* side, float dt,
                       float half_dt, const float* dt_device,
                       double* position_hi_out, float* position_out,
                       int32_t position_stride, float* momentum_out,
                       const float* reference_position, float threshold_sq,
                       int32_t* gate, void* stream);

/* jmd_nve_kick_drift in the x64 form (simulate.py:227-239 with float64
 * state): position, momentum in place (float64 [n, dim]), force float64,
 * mass NULL (mass_scalar) or device float64 [n]; dt and half_dt the float32
 * values the reference casts them to.  position_f32 receives the float32
 * mirror of the new positions (row stride position_stride) from which lists
 * are built, and the skin criterion is evaluated on it as in
 * jmd_nve_kick_drift. */
int jmd_nve_kick_drift_f64(double* position, double* momentum, const double* force,
                           const double* mass, double mass_scalar, int64_t n, int32_t dim,
                           int32_t periodic, const float* side, double dt, double half_dt,
                           float* position_f32, int32_t position_stride,
                           const float* reference_position, float threshold_sq,
                           int32_t* gate, void* stream);

/* Nose-Hoover chain thermostat of simulate.nvt_nose_hoover
 * (simulate.py:328-347,384-672): the chain (chain = float64 [3, L]: positions,
 * momenta, masses) is advanced on the device by one thread and hands the
 * particles one momentum factor per half step (momentum_scale, device
 * float64 [1]).  mode bits: 4 (init) kinetic = sum of ke_partials, chain at
 * rest; 1 (finish a step) kinetic = sum of ke_partials (after the closing
 * kick), half step; 2 (start a step) chain masses kT tau^2 (x dof for the
 * first), half step.  dt, kT, tau are the float32 values the reference
 * uses; dof = n * dim. */
int jmd_nhc_step(int32_t mode, double* chain, int32_t chain_length, int32_t chain_steps,
                 int32_t sy_steps, float dt, float kT, float tau, double dof,
                 double* kinetic, const double* ke_partials, int32_t n_partials,
                 double* momentum_scale, void* stream);

/* p += half_dt F (force may be NULL: no kick) and per-block float64 partials
 * of the kinetic energy sum p^2 / (2 m) (jmd_kick_kinetic_partials(n) of
 * them) in a fixed order. */
int jmd_kick_kinetic_partials(int64_t n);
int jmd_kick_kinetic(float* momentum, const float* force, const float* mass, float mass_scalar,
                     int64_t n, int32_t dim, float half_dt, double* ke_partials, void* stream);

/* jmd_kick_kinetic followed by jmd_nhc_step(mode, ...) in one launch, with
 * the same results bit for bit: the closing kick and kinetic energy of
 * nvt_nose_hoover's apply_fn and the chain half steps around the step
 * boundary (simulate.py:492-515,650-669).  ke_partials holds
 * jmd_kick_kinetic_partials(n) entries; ticket is a device uint32 that is 0
 * before the first call and is left at 0 by every call.  Calls that share a
 * ticket must not overlap. */
int jmd_kick_kinetic_nhc(float* momentum, const float* force, const float* mass,
                         float mass_scalar, int64_t n, int32_t dim, float half_dt,
                         double* ke_partials, uint32_t* ticket, int32_t mode, double* chain,
                         int32_t chain_length, int32_t chain_steps, int32_t sy_steps, float dt,
                         float kT, float tau, double dof, double* kinetic,
                         double* momentum_scale, void* stream);

/* p *= momentum_scale[0] (the chain's closing factor at the end of a block). */
int jmd_scale_momentum(float* momentum, int64_t n, int32_t dim, const double* scale,
                       void* stream);

/* jmd_nve_kick_drift with the thermostat factor applied to p before the
 * kick (p' = s p + half_dt F), as nvt_nose_hoover's apply_fn scales the
 * momenta before velocity_verlet (simulate.py:644-652). */
int jmd_nvt_kick_drift(const double* position_hi, const float* momentum, const float* force,
                       const float* mass, float mass_scalar, int64_t n, int32_t dim,
                       int32_t periodic, const float* side, float dt, float half_dt,
                       double* position_hi_out, float* position_out, int32_t position_stride,
                       float* momentum_out, const float* reference_position,
                       float threshold_sq, int32_t* gate, const double* momentum_scale,
                       void* stream);

/* The force pass of simulate.velocity_verlet fused with the start of the
 * next step (jax_md/simulate.py:227-243, partition.py:1141-1154, the loop of
 * examples/nve_neighbor_list.py:186-201): over a Dense list,
 *   F = F(R) (jmd_pair_energy_force, written to force), p += half_dt F;
 * then, per particle, the skin criterion of R against reference_position
 * raises *gate_next (the next step's update decision, as
 * jmd_neighbor_list_check; gate_next may be NULL), and with drift != 0 the
 * next step's opening half: p += half_dt F, R' = shift(R, dt p / m) as in
 * jmd_nve_kick_drift (position_hi in place, float32 R' to position_out, a
 * second buffer of R's layout that must not alias position).  momentum holds
 * the last momentum written.  The shift's space is shift_periodic /
 * shift_side; dt_device (may be NULL) overrides dt with half_dt = float32(dt
 * / 2).  No counterpart in the reference as a single call: with the list
 * update gated on gate_next it leaves one pass over the particles per step
 * outside the force launch. */
int jmd_nve_force_kick_drift(const float* position, int64_t n, int32_t dim,
                             int32_t position_stride, int32_t periodic, const float* side,
                             const jmd_pair_potential* potential, const int32_t* idx,
                             int64_t idx_width, float* force, float* momentum,
                             float half_dt, const float* dt_device, double* position_hi,
                             const float* mass, float mass_scalar, int32_t shift_periodic,
                             const float* shift_side, float dt,
                             const float* reference_position, float threshold_sq,
                             int32_t* gate_next, int32_t drift, float* position_out,
                             void* stream);

/* One step's neighbour-list update fused with the opening half of the next
 * velocity-Verlet step: the loop body
 *   nbrs = nbrs.update(R); state = apply_fn(state, neighbor=nbrs)
 * of the reference's NVE driver (examples/nve_neighbor_list.py:186-201,
 * partition.py:1141-1164, simulate.py:227-243) up to the force evaluation.
 * The list part is jmd_neighbor_list_build with `gate` (rebuild only when the
 * gate is raised; the gate is cleared by the rebuild).  In the same pass over
 * the particles as the binning, p += half_dt F and R' = shift(R, dt p / m) as
 * in jmd_nve_kick_drift (position_hi and momentum updated in place, shift in
 * the space given by shift_periodic / shift_side), with R' written to
 * position_out, a second [n, dim] buffer of the same layout that must not
 * alias position (the list stages still read R), and the skin criterion of
 * R' against the list's reference positions (R itself after a rebuild)
 * raising gate_next (a second word, distinct from gate).  The caller then
 * evaluates the force at position_out.  No counterpart in the reference as a
 * single call: it saves one kernel launch per step. */
int jmd_nve_list_update_kick_drift(
    const float* position, int64_t n, int32_t dim, int32_t position_stride,
    int32_t periodic, const float* side, float cutoff_sq, int32_t use_cell_list,
    const int32_t* cells_per_side, const float* cell_size, int32_t cell_capacity,
    int32_t format, int32_t mask_self, int32_t* idx, int64_t max_occupancy,
    int32_t* occupancy, uint8_t* error, float* reference_position, int32_t* gate,
    void* workspace, double* position_hi, float* momentum, const float* force,
    const float* mass, float mass_scalar, int32_t shift_periodic,
    const float* shift_side, float dt, float half_dt, const float* dt_device,
    float* position_out, float threshold_sq, int32_t* gate_next, void* stream);

/* FIRE bookkeeping after the NVE step of minimize.fire_descent.apply_fn
 * (jax_md/minimize.py:184-224).  scalars (device float64 [3]) receives
 * sum F^2, sum P^2 and sum F.P (float64 sums).  Then the one-block update
 * applies, in the reference's order and with float16 dt / alpha arithmetic
 * (util.static_cast, util.py:58-64):
 *   P += alpha (F |P| / |F| - P);  n_pos, dt, alpha update;  P *= (F.P >= 0)
 * fire_state (device): float32 [dt, alpha] holding float16-representable
 * values, int32 n_pos at fire_n_pos.  params (host float32 [7]):
 * dt_start, dt_max, n_min, f_inc, f_dec, alpha_start, f_alpha, already
 * rounded to their static_cast dtypes. */
int jmd_fire_update(float* momentum, const float* force, int64_t n,
                    int32_t dim, float* fire_state, int32_t* fire_n_pos,
                    const float* params, double* scalars, void* workspace,
                    void* stream);

/* Scratch bytes for jmd_fire_update. */
size_t jmd_fire_workspace_bytes(int64_t n);

/* Multi-image neighbour list of a small periodic cell
 * (custom_partition.neighbor_list_multi_image's build, custom_partition.py:
 * 212-626): every pair (i, j) against every lattice shift s in
 * [-shift_max[d], shift_max[d]] (meshgrid order, first dimension slowest),
 * d = (x_j + s T^T) - x_i in real space with x = T u when fractional (box =
 * the [dim, dim] matrix T, host, row-major), kept when d^2 < cutoff_sq and
 * not (i = j, s = 0).  format JMD_SPARSE / JMD_ORDERED_SPARSE: idx [2,
 * capacity] (receiver i, sender j) in (s, i, j) order, edge_shift [capacity,
 * dim] int32; OrderedSparse keeps i < j at s = 0 and canonical shifts (first
 * nonzero component positive) otherwise.  JMD_DENSE: idx [n, capacity], per
 * atom the capacity nearest kept (s, j) by d^2 (ties by s n + j),
 * edge_shift [n, capacity, dim].  Padding: index n, shift 0.  occupancy
 * (device uint64 [1]) receives the total kept (Sparse) or the largest per-atom
 * count (Dense); overflow (device uint8 [1]) = occupancy > capacity.  With
 * capacity 0 or idx NULL only occupancy is produced (the allocation probe).
 * workspace: jmd_multi_image_workspace_bytes(n, stencil size). */
size_t jmd_multi_image_workspace_bytes(int64_t n, int32_t n_shift);

/* max_i |x_i - x_ref,i|^2 in real space, unwrapped (x = T u, x_ref = T_ref
 * u_ref when fractional), into out (device float [1]): the displacement of
 * neighbor_list_multi_image's rebuild test (custom_partition.py:961-985). */
int jmd_multi_image_max_displacement_sq(const float* position, const float* reference,
                                        int64_t n, int32_t dim, int32_t fractional,
                                        const float* box, const float* reference_box,
                                        float* out, void* stream);
int jmd_multi_image_build(const float* position, int64_t n, int32_t dim, int32_t fractional,
                          const float* box, const int32_t* shift_max, float cutoff_sq,
                          int32_t format, int32_t* idx, int32_t* edge_shift, int64_t capacity,
                          unsigned long long* occupancy, uint8_t* overflow, void* workspace,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* JAXMD_NATIVE_H_ */
