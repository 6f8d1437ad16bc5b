/*
 * cusoap_amd -- C ABI of the MI355X (gfx950) SOAP descriptor kernels.
 *
 * The reference (TACC/cuSOAP) has no C/FFI boundary: its hot path is the
 * Python class `SOAP` in src/cusoap/soap_generator.py.  These entry points
 * replace the private stages of that class; the Python mirror
 * `cusoap_amd.SOAP` keeps the public signatures and binds these functions
 * with ctypes.  Each declaration names the reference code it replaces.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers allocated by the caller
 *    (caller-owned memory, typically PyTorch's caching allocator), except
 *    the `*_host` out-parameters and the plan/basis structs, which live in
 *    host memory.
 *  - `stream` is a hipStream_t passed as void*; all work is enqueued on it.
 *    Functions marked "synchronises" wait for the stream before returning.
 *  - Every function returns CUSOAP_OK or an error code; the text of the last
 *    error of the calling thread is returned by cusoap_last_error().
 *  - Real spherical harmonics use the l-major, m = -l..l order with positive
 *    real harmonics (Y_1,-1..1 proportional to y, z, x).
 *  - Coefficient arrays are c[center][species][n][lm], lm = l*l + l + m.
 */
#ifndef CUSOAP_AMD_H
#define CUSOAP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CUSOAP_OK = 0,
    CUSOAP_ERR_ARG = 1,          /* invalid argument (maps to ValueError)              */
    CUSOAP_ERR_HIP = 2,          /* HIP runtime error (RuntimeError)                   */
    CUSOAP_ERR_UNSUPPORTED = 3,  /* parameter outside the compiled range (NotImplemented) */
    CUSOAP_ERR_CAPACITY = 4,     /* a per-center neighbour count exceeds a kernel limit  */
    CUSOAP_ERR_WORKSPACE = 5     /* workspace smaller than the size query asked for      */
};

enum { CUSOAP_RBF_GTO = 0, CUSOAP_RBF_POLY = 1 };
enum { CUSOAP_W_NONE = 0, CUSOAP_W_POLY = 1, CUSOAP_W_POW = 2, CUSOAP_W_EXP = 3 };
/* Arithmetic of the coefficient and adjoint stages.  F64: everything in
 * double.  MIXED: radial functions (exponentials and the Loewdin beta
 * contraction) in double per edge, harmonics and accumulation in float.
 * F32: everything in float (the Loewdin contraction then loses digits at
 * large n_max). */
enum { CUSOAP_PREC_F64 = 0, CUSOAP_PREC_F32 = 1, CUSOAP_PREC_MIXED = 2 };
enum { CUSOAP_DT_F64 = 0, CUSOAP_DT_F32 = 1 };
enum { CUSOAP_MODE_PS = 0, CUSOAP_MODE_CC = 1 };
enum { CUSOAP_MFMA_F16 = 0, CUSOAP_MFMA_BF16 = 1 };

/* Radial basis tables, all device pointers to float64.  Mirrors the state the
 * reference builds in SOAP.__init__ (soap_generator.py:600-634, 690-819). */
typedef struct cusoap_basis {
    int32_t n_species, n_max, l_max, rbf;
    double eta;                 /* 1 / (2 sigma^2)                                  */
    double cutoff;              /* r_cut + sigma sqrt(-2 ln 1e-3)                   */
    int32_t weighting;          /* CUSOAP_W_*                                       */
    int32_t quad_n;             /* polynomial basis: Gauss-Legendre points          */
    double w_r0, w_c, w_d, w_m, w0;
    const double* gto_gamma;    /* (l_max+1, n_max)  alpha eta / (alpha + eta)      */
    const double* gto_pref;     /* (l_max+1, n_max)  pi^1.5 (eta/p)^l p^-1.5        */
    const double* gto_beta;     /* (l_max+1, n_max, n_max)  Loewdin S_l^-1/2        */
    const double* poly_rx;      /* (quad_n)          quadrature nodes on [0, r_cut] */
    const double* poly_rxw;     /* (quad_n)          4 pi rx^2 w_q                  */
    const double* poly_gss;     /* (n_max, quad_n)   orthonormal g_n(rx_q)          */
    const double* poly_B0;      /* (n_max, l_max+1)  r -> 0 limit of B              */
    const double* poly_Bh0;     /* (n_max, l_max+1)  r -> 0 limit of Bhat           */
    const double* self_l0;      /* (n_max)           l=0 self coefficient (incl. w0, Y00) */
    /* Tabulated radial model (the learned radial surrogate, see cusoap_radial_table),
     * float32 device array or NULL for the exact radial functions.  Layout
     * ((S+1) S slots, rad_intervals+1 nodes, l_max+1, n_max, 2): the pair
     * (B_nl(r_i), dB_nl/dr(r_i)) at r_i = i rad_rmax / rad_intervals; slot = center_slot * S +
     * neighbour species, center_slot = the center atom's species, or S for a
     * center that is not an atom.  Kernels interpolate cubic Hermite in r. */
    const float* rad_table;
    int32_t rad_intervals;
    double rad_rmax;
    /* Chebyshev radial model (radial="chebyshev", and the learned model in that form),
     * float32 device array or NULL.  Two intervals [0, cheb_split] and [cheb_split,
     * cutoff], cheb_terms (32) terms each.  Layout ((S+1) S slots, 2, l_max+1, 2
     * intervals, cheb_terms, 8): the coefficients a_j of f_nl(r) = r^l B_nl(r) (index 0)
     * and of df_nl/dr (index 1) in T_j(t), t = 2 (r - lo) / (hi - lo) - 1 on the interval
     * [lo, hi] holding r, n padded to 8, slots as for rad_table.  When set, the float
     * coefficient stage (cusoap_coefficients with a float precision) runs on the matrix
     * cores (fp16 hi/lo operands, float32 accumulation) and does not read rad_table; the
     * adjoint edge pass reads rad_table (the same radial model tabulated). */
    const float* rad_cheb;
    int32_t cheb_terms;
    double cheb_split;
    /* The same value coefficients as fp16 matrix-core operands: ((S+1) S slots, l_max+1,
     * 2 intervals, 64 lanes, 16) halves, lane = n + 16 g holding 8 hi then 8 lo halves of
     * a_l[n][k] 2^e_ln at k = 4 g + (j % 4) + 16 (j / 4), and cheb_rowinv ((S+1) S, l_max+1,
     * 8) float32 = 2^-e_ln.  Required with rad_cheb. */
    const void* cheb_frag;
    const float* cheb_rowinv;
} cusoap_basis_t;

/* Cell-list geometry of one neighbour search. */
typedef struct cusoap_nl_plan {
    double origin[3];
    double cell[3];
    int32_t dims[3];
    int32_t n_cells;
    int64_t n_ext;              /* n_images * n_atoms                                */
    size_t ws_bytes;            /* workspace needed by nl_count / nl_fill            */
} cusoap_nl_plan_t;

/* Library information. */
const char* cusoap_last_error(void);
int cusoap_version(void);
int cusoap_max_l(void);
int cusoap_max_n(void);

/* Kernel timing: when enabled, the library records HIP events around its main
 * kernels on the launch stream ("nl_count", "nl_fill", "coefficients",
 * "power_spectrum", "cc_projection", "jacobian_edges", "jacobian_dense",
 * "vjp_adjoint", "vjp_edges").  cusoap_profile_get synchronises on the recorded
 * events, returns the summed milliseconds and launch count of `name`, and
 * clears them.  No reference counterpart: it backs create_with_profile
 * (soap_generator.py:1815-1868) and the benchmark. */
void cusoap_profile_enable(int32_t on);
int cusoap_profile_get(const char* name, double* total_ms, int32_t* count);

/* ---- neighbour search (replaces _build_neighbor_list / _radius_edges /
 *      _extend_periodic, soap_generator.py:1015-1181) ---------------------
 * Ext index of a neighbour = image * n_atoms + atom, images in the meshgrid-'ij'
 * order of the reference's _pbc_offsets (soap_generator.py:382-396); shifts
 * are (n_img, 3) Cartesian image translations.  Edges of a center are sorted by
 * ext index, centers ascending: the order of the reference's `nonzero`.  The
 * distance test is d^2 = ((-2 c.n) + |c|^2) + |n|^2 <= cutoff^2 in float64
 * with fixed summation order and no fused multiply-add.
 * drop_self != 0 removes edges with |r| <= 1e-12 (the create() path).
 */
/* Bounding box of the centers and cell grid; synchronises.  `tmp` >= 16 KiB device. */
int cusoap_nl_plan(const double* centers, int32_t n_centers, int32_t n_atoms, int32_t n_img,
                   double cutoff, void* tmp, cusoap_nl_plan_t* plan, void* stream);
/* Bins the images into the cell grid and counts edges per center; writes
 * offsets (n_centers+1, int64, exclusive scan).  Synchronises and returns the
 * edge total and the largest per-center count. */
int cusoap_nl_count(const cusoap_nl_plan_t* plan, const double* positions, int32_t n_atoms,
                    const double* shifts, int32_t n_img, const double* centers, int32_t n_centers,
                    double cutoff, int32_t drop_self, void* ws, size_t ws_bytes,
                    int64_t* offsets, int64_t* total_host, int32_t* max_count_host, void* stream);
/* Writes the sorted ext indices (int32) of every center's edges.  `ws` must be the
 * workspace that cusoap_nl_count filled for the same plan and centers (it holds the
 * binned images and the rows the count pass already sorted). */
int cusoap_nl_fill(const cusoap_nl_plan_t* plan, const double* positions, int32_t n_atoms,
                   const double* shifts, int32_t n_img, const double* centers, int32_t n_centers,
                   double cutoff, int32_t drop_self, void* ws, size_t ws_bytes,
                   const int64_t* offsets, int32_t max_count, int32_t* nbr, void* stream);
/* plan + count + fill in one call, for callers that can size the buffers ahead (from
 * an earlier call on a similar system): no host work between the three stages.
 * Returns CUSOAP_ERR_WORKSPACE when ws_bytes < plan->ws_bytes (plan filled, nothing
 * counted: call again with that workspace, or continue with cusoap_nl_count).  When
 * the edge total exceeds nbr_capacity the fill is skipped (the return is CUSOAP_OK
 * with *total_host > nbr_capacity): complete it with cusoap_nl_fill on an nbr of
 * *total_host entries and the same workspace. */
int cusoap_nl_build(const double* positions, int32_t n_atoms, const double* shifts, int32_t n_img,
                    const double* centers, int32_t n_centers, double cutoff, int32_t drop_self, void* tmp,
                    void* ws, size_t ws_bytes, int64_t* offsets, int32_t* nbr, int64_t nbr_capacity,
                    cusoap_nl_plan_t* plan, int64_t* total_host, int32_t* max_count_host, void* stream);

/* ---- density coefficients (replaces _coefficients_gto / _coefficients_polynomial
 *      / _add_self_terms, soap_generator.py:1294-1435, 821-850, and the
 *      Cacc accumulation of derivatives_analytical_ps, 2414-2423) ----------
 * One fused pass per center over its edges: real solid harmonics, radial
 * functions (with weighting), accumulation and the Loewdin contraction.
 * create_path != 0: edges with |r| <= 1e-12 are skipped and self_l0 is added
 * for centers with center_self_species >= 0 (index centers); otherwise every
 * edge is kept (the closed-form derivative path).  center_self_species (the
 * species of the center atom, -1 for a coordinate center) also selects the
 * radial-table slot when basis->rad_table is set; it is then required.
 * coeffs: (n_centers, n_species, n_max, (l_max+1)^2), float64 for PREC_F64,
 * float32 for PREC_MIXED / PREC_F32. */
int cusoap_coefficients(const cusoap_basis_t* basis, int32_t precision,
                        const double* positions, int32_t n_atoms, const int32_t* atom_species,
                        const double* shifts, int32_t n_img,
                        const double* centers, int32_t n_centers, const int32_t* center_self_species,
                        const int64_t* offsets, const int32_t* nbr, int32_t create_path,
                        void* coeffs, void* stream);

/* ---- descriptor assembly (replaces _power_spectrum / _projection_cc /
 *      _to_dscribe_cc, soap_generator.py:1599-1689) --------------------------
 * feat_table: (n_feat) packed uint32 j | jd<<8 | n<<16 | nd<<22 | l<<28 in the
 * DScribe order of _build_feature_slices (soap_generator.py:1215-1291). */
int cusoap_power_spectrum(const cusoap_basis_t* basis, const void* coeffs, int32_t coeffs_dtype,
                          int32_t n_centers, const uint32_t* feat_table, int32_t n_feat,
                          void* out, int32_t out_dtype, void* stream);
int cusoap_cc_projection(const cusoap_basis_t* basis, const void* coeffs, int32_t coeffs_dtype,
                         int32_t n_centers, void* out, int32_t out_dtype, void* stream);
/* The same power spectrum on the matrix cores: per center and l the Gram matrix
 * of the (S n) x (2l+1) coefficient block from 16x16x16 MFMA tiles with fp16
 * (kind = CUSOAP_MFMA_F16) or bf16 (CUSOAP_MFMA_BF16) operands and float32
 * accumulation; split != 0 carries each operand as hi + lo and sums three
 * products (hi hi + hi lo + lo hi).  The block is scaled by a power of two before
 * rounding.  pair_feature: (l_max+1, S n, S n) int32 feature index of each
 * channel pair or -1, symmetric (as for cusoap_vjp).  No reference counterpart
 * for the arithmetic; it replaces the einsum of _power_spectrum
 * (soap_generator.py:1634-1689). */
int cusoap_power_spectrum_mfma(const cusoap_basis_t* basis, const void* coeffs, int32_t coeffs_dtype,
                               int32_t n_centers, const int32_t* pair_feature, int32_t n_feat,
                               int32_t kind, int32_t split, void* out, int32_t out_dtype, void* stream);

/* ---- dense analytical Jacobian (replaces derivatives_analytical_ps and its
 *      Triton kernels _ps_uv_kernel / _ps_jacobian_dense_kernel /
 *      _cc_jacobian_dense_kernel, soap_generator.py:50-220, 2270-2626, and the
 *      autograd path derivatives_analytical, 2007-2113, for periodic cells,
 *      weighting and attach) ----------------------------------------------------
 * out: (n_centers, n_inc, 3, n_feat) in out_dtype.  coeffs: float64, from
 * cusoap_coefficients on the same edges (create_path = 0 for the closed-form
 * convention).  inv_atom: (n_atoms) int32, output row of each atom or -1.
 * center_attach: (n_centers) int32 atom an index center moves with, or -1, or
 * NULL (fixed centers, the reference default).  Edges must keep self edges.
 * compute_dtype: arithmetic of rows fed by a single edge (CUSOAP_DT_F64 as the
 * reference; CUSOAP_DT_F32 only with float32 output); records, rows with
 * several edges and the attach row are always float64. */
int cusoap_jacobian_workspace_size(const cusoap_basis_t* basis, int32_t mode, int64_t n_edges,
                                   int32_t n_centers, int32_t n_inc, size_t* bytes);
int cusoap_jacobian_dense(const cusoap_basis_t* basis, int32_t mode,
                          const double* positions, int32_t n_atoms, const int32_t* atom_species,
                          const double* shifts, int32_t n_img,
                          const double* centers, int32_t n_centers, const int32_t* center_attach,
                          const int64_t* offsets, const int32_t* nbr, int64_t n_edges,
                          const double* coeffs, const int32_t* inv_atom, int32_t n_inc,
                          const uint32_t* feat_table, int32_t n_feat,
                          void* ws, size_t ws_bytes, void* out, int32_t out_dtype,
                          int32_t compute_dtype, void* stream);

/* ---- adjoint: vector-Jacobian product (no reference counterpart; the
 *      reference contracts its dense Jacobian, soap_generator.py:2270-2626) ---
 * grad_pos (n_atoms, 3) float64 receives sum_{c,f} grad_desc[c,f] d desc[c,f] / d x
 * (it is zeroed first).  grad_desc: (n_centers, n_feat) in gd_dtype, or with
 * gd_broadcast != 0 a single (n_feat) row used for every center.  coeffs:
 * (n_centers, S, n, L2) in coeffs_dtype, from cusoap_coefficients on the same
 * centers.  precision: CUSOAP_PREC_* of the edge pass (the polynomial basis
 * always runs in float64).  pair_feature: (l_max+1, S n, S n) int32 feature
 * index of each channel pair or -1 (power spectrum only).  center_attach as
 * for the Jacobian.  center_species: (n_centers) species of the center atom or
 * -1, required when basis->rad_table is set (table slot), else may be NULL.
 * Forces are accumulated with float64 atomics (the summation order, and so the
 * last bits, vary from run to run); cusoap_vjp_ordered below sums in a fixed
 * order.  The workspace also holds the centers' Morton order (a sort of their
 * coordinates, run on a second stream of the library's own beside the first
 * kernel): the edge pass takes groups of centers along it, so its speed does
 * not depend on the order of the caller's atoms (cusoap_vjp_noalloc keeps the
 * given order). */
int cusoap_vjp_workspace_size(const cusoap_basis_t* basis, int32_t n_centers, size_t* bytes);
int cusoap_vjp(const cusoap_basis_t* basis, int32_t mode, int32_t precision,
               const double* positions, int32_t n_atoms, const int32_t* atom_species,
               const double* shifts, int32_t n_img,
               const double* centers, int32_t n_centers, const int32_t* center_attach,
               const int32_t* center_species, const int64_t* offsets, const int32_t* nbr,
               const void* coeffs, int32_t coeffs_dtype, const int32_t* pair_feature,
               const void* grad_desc, int32_t gd_dtype, int32_t gd_broadcast, int32_t n_feat,
               void* ws, size_t ws_bytes, double* grad_pos, void* stream);

/* ---- fixed-order force sums (no reference counterpart: the reference's
 *      index_add_ scatters, soap_generator.py:1375-1381, 2422, are atomic and
 *      run-to-run variable on the GPU) -------------------------------------------
 * cusoap_edge_reverse: rev (offsets[n_centers]) int32, for each edge c -> a the
 * position of the reverse edge a -> c (image with the opposite shift) in atom
 * a's range of the same lists, or -1 when it is missing (a pair within rounding
 * of the cutoff) or the lists are not the atoms' own (centers[c] or centers[a]
 * differs from the atom's position, bitwise).  Depends on the neighbour lists
 * only, so it may run on another stream while the coefficients are formed.
 * Workspace: cusoap_edge_reverse_workspace_size bytes (a 64-byte search block
 * per center).
 * cusoap_vjp_ordered: cusoap_vjp for every atom its own center (n_centers ==
 * n_atoms, centers[c] == positions[c]) with rev from cusoap_edge_reverse and
 * n_edges = offsets[n_centers]: each edge's contribution is stored at its own
 * index, and each atom's force is the float64 sum, in a fixed order, of the
 * contributions of its reverse edges, so the result is the same bits on every
 * run.  Edges with rev = -1 add atomically (as cusoap_vjp).  Workspace:
 * cusoap_vjp_ordered_workspace_size bytes. */
int cusoap_edge_reverse_workspace_size(int32_t n_centers, size_t* bytes);
int cusoap_edge_reverse(const double* positions, int32_t n_atoms, const double* shifts, int32_t n_img,
                        const double* centers, int32_t n_centers, const int64_t* offsets, const int32_t* nbr,
                        void* ws, size_t ws_bytes, int32_t* rev, void* stream);
int cusoap_vjp_ordered_workspace_size(const cusoap_basis_t* basis, int32_t precision, int32_t n_centers,
                                      int64_t n_edges, size_t* bytes);
int cusoap_vjp_ordered(const cusoap_basis_t* basis, int32_t mode, int32_t precision,
                       const double* positions, int32_t n_atoms, const int32_t* atom_species,
                       const double* shifts, int32_t n_img,
                       const double* centers, int32_t n_centers, const int32_t* center_attach,
                       const int32_t* center_species, const int64_t* offsets, const int32_t* nbr,
                       const int32_t* rev, int64_t n_edges,
                       const void* coeffs, int32_t coeffs_dtype, const int32_t* pair_feature,
                       const void* grad_desc, int32_t gd_dtype, int32_t gd_broadcast, int32_t n_feat,
                       void* ws, size_t ws_bytes, double* grad_pos, void* stream);

/* Allocation-free cusoap_vjp.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving or MD loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as cusoap_vjp without the
 * workspace: the dL/dc rows live in a static device array of
 * CUSOAP_VJP_NOALLOC_SCRATCH_BYTES (64 MiB) that every GPU the library runs on holds until
 * the process ends, and the centers are processed in as many chunks as that array needs
 * (forces add up over the chunks).  Calls no hipMalloc, hipFree or allocating library
 * routine, and zeroes grad_pos with a kernel (no memset node).  All calls on one GPU share
 * its array: calls that overlap in time on one GPU (other streams, other threads) race.
 * Returns CUSOAP_ERR_CAPACITY when a single center's rows do not fit the array. */
#define CUSOAP_VJP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int cusoap_vjp_noalloc(const cusoap_basis_t* basis, int32_t mode, int32_t precision,
                       const double* positions, int32_t n_atoms, const int32_t* atom_species,
                       const double* shifts, int32_t n_img,
                       const double* centers, int32_t n_centers, const int32_t* center_attach,
                       const int32_t* center_species, const int64_t* offsets, const int32_t* nbr,
                       const void* coeffs, int32_t coeffs_dtype, const int32_t* pair_feature,
                       const void* grad_desc, int32_t gd_dtype, int32_t gd_broadcast, int32_t n_feat,
                       double* grad_pos, void* stream);

/* ---- Loewdin orthonormalisation S^-1/2 in float64 (replaces
 *      _lowdin_invsqrt, soap_generator.py:339-343, which calls torch.linalg.eigh
 *      on the device) ----------------------------------------------------------
 * S, out: (batch, n, n) row-major symmetric; cyclic Jacobi, eigenvalues clamped
 * at eps before the inverse square root.  n <= 32. */
int cusoap_lowdin_invsqrt_f64(const double* S, int32_t batch, int32_t n, double eps,
                              double* out, void* stream);

/* ---- learned radial surrogate (no reference counterpart; it stands in for the
 *      exact radial integral of _coefficients_gto and derivatives_analytical_ps,
 *      soap_generator.py:1358-1391, 2389-2405, where its held-out error allows) --
 * One network per species pair (center species, neighbour species), float32:
 * x in [-1, 1] -> 64 tanh -> 64 tanh -> n_out linear.  Parameters per pair, in
 * this order: W1[64] b1[64] W2[64][64] b2[64] W3[n_out][64] b3[n_out]; pairs are
 * contiguous.  n_hidden must be 64 and n_out <= 128. */
/* Number of float parameters of one network, or -1 for an unsupported shape. */
int64_t cusoap_mlp_param_count(int32_t n_hidden, int32_t n_out);
/* Exact float64 orthonormal GTO radial functions at n_pts radii:
 * B[p][l][n] = sum_k beta_l[n][k] pref_kl exp(-gamma_kl r_p^2) and dBdr = dB/dr. */
int cusoap_radial_exact(const cusoap_basis_t* basis, const double* r, int32_t n_pts,
                        double* B, double* dBdr, void* stream);
/* Network outputs and input derivatives of every pair at n_pts inputs x:
 * y, dydx (n_pairs, n_pts, n_out) float32. */
int cusoap_mlp_eval(const float* params, int32_t n_pairs, int32_t n_hidden, int32_t n_out,
                    const float* x, int32_t n_pts, float* y, float* dydx, void* stream);
/* Loss and gradient over a minibatch of `batch` samples drawn (with replacement,
 * by a hash of seed, step and position) from n_samples training points:
 * L = 1/(n_out batch) sum [w_val_o (y - t)^2 + w_der_o (dy/dx - dt)^2].
 * x (n_samples); t, dt (n_pairs, ...) with target_pair_stride elements between
 * pairs (0 = shared targets), each (n_samples, n_out).  grad (n_pairs, P),
 * loss (n_pairs) or NULL.  Deterministic (fixed-order reduction). */
int cusoap_mlp_train_workspace_size(int32_t n_pairs, int32_t n_hidden, int32_t n_out, int32_t batch,
                                    size_t* bytes);
int cusoap_mlp_loss_grad(const float* params, int32_t n_pairs, int32_t n_hidden, int32_t n_out,
                         const float* x, const float* t, const float* dt, int32_t n_samples,
                         int64_t target_pair_stride, const float* w_val, const float* w_der,
                         int32_t batch, uint32_t seed, int64_t step, void* ws, size_t ws_bytes,
                         float* grad, float* loss, void* stream);
/* Adam (Kingma & Ba 2015) with bias correction, step counted from 1. */
int cusoap_adam_step(float* params, const float* grad, float* m, float* v, int64_t n, float lr,
                     float beta1, float beta2, float eps, int64_t step, void* stream);
/* Exact least-squares refit of the output layer (W3, b3) of every pair for the
 * current hidden layers, minimising the same loss over all n_samples points
 * (float64 normal equations, Cholesky; ridge scales trace(A)/65 on the diagonal). */
int cusoap_mlp_refit_workspace_size(int32_t n_pairs, int32_t n_hidden, int32_t n_out, int32_t n_samples,
                                    size_t* bytes);
int cusoap_mlp_refit_output(float* params, int32_t n_pairs, int32_t n_hidden, int32_t n_out,
                            const float* x, const float* t, const float* dt, int32_t n_samples,
                            int64_t target_pair_stride, const float* w_der, double ridge,
                            void* ws, size_t ws_bytes, void* stream);
/* Radial table for basis->rad_table (layout there) on n_intervals intervals of
 * [0, r_max].  Node values come from the network of pair p = center species * S +
 * neighbour species where use_pair[p] != 0 and 0 < r_i <= r_fit, and from the
 * exact float64 function elsewhere, including r_i = 0 (and for every slot when
 * params is NULL).  The network models g_nl = u^l B_nl, u = r / r_fit (input
 * x = 2u - 1, outputs l-major j = l n_max + n, g = scale[p][j] y, dg/dr =
 * scale[p][j] (2 / r_fit) dy/dx); the table holds B = g / u^l and dB/dr =
 * g'/u^l - l g / (r_fit u^(l+1)).  dB/dr is 0 at r = 0. */
int cusoap_radial_table_workspace_size(const cusoap_basis_t* basis, int32_t n_intervals, size_t* bytes);
int cusoap_radial_table(const cusoap_basis_t* basis, const float* params, int32_t n_hidden,
                        const int32_t* use_pair, const float* scale, double r_fit,
                        int32_t n_intervals, double r_max, void* ws, size_t ws_bytes,
                        float* table, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CUSOAP_AMD_H */
