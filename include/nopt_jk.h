/* RI-J/K Fock build on one MI355X (gfx950): the C-ABI boundary.
 *
 * This replaces NOPT's RI layer (include/RI.h:6-23, src/RI.cpp) for the
 * Fock build. NOPT keeps module globals (RI_AA, AUX_Vm05, RI_CA, RI_CA2,
 * src/RI.cpp:16-32); here that state lives in a context that owns the B
 * tensor resident in HBM.
 *
 * Conventions
 *  - Every matrix is row-major FP64 on the host side unless stated.
 *  - Densities D are normalised to 1 as in NOPT: D = sum_i C_i C_i^T over the
 *    occupied rows of the orbital matrix (gen_HF_DM, src/SCF.cpp:63-73); the
 *    closed-shell Fock matrix is F = H + 2J - K (molecule::calc_F_AO,
 *    src/molecule.cpp:4407-4438).
 *  - Orbital factors are passed as NOPT passes MO: rows are orbitals,
 *    C[i*nbf + m], and K is built from the first nocc rows
 *    (DM_to_F_transform_RI, src/RI.cpp:146-171).
 *  - Return value: 0 (NOPT_OK) or a NOPT_E* code; nopt_last_error() gives the
 *    message. NOPT itself exits or throws on these conditions
 *    (src/RI.cpp:36-39, std::bad_alloc); here they are reported.
 *  - Calls are host-synchronous unless they take a stream; one context must
 *    not be used from two threads at once.
 *  - Work enqueued on one context runs in call order whatever the streams:
 *    every call waits, on its own stream, for the context's previous work
 *    (an event recorded at the end of each call), because all builds of a
 *    context share its work buffers. Builds of different contexts on
 *    different streams may overlap.
 */
#ifndef NOPT_JK_H
#define NOPT_JK_H
#ifdef __cplusplus
extern "C" {
#endif

#define NOPT_OK 0
#define NOPT_EINVAL 1 /* bad argument or shape */
#define NOPT_ENOMEM 2 /* device or host allocation failed */
#define NOPT_ECHOL 3  /* Coulomb metric not positive definite */
#define NOPT_EHIP 4   /* HIP runtime error */
#define NOPT_EIO 5    /* file could not be read */
#define NOPT_ESTATE 6 /* call out of order (e.g. a build before set_B) */

/* Arithmetic of the half-transform and exchange contraction. J, the metric
 * and the Fock assembly are FP64 in every mode. */
#define NOPT_PREC_FP64 0   /* FP64 MFMA */
#define NOPT_PREC_FP32 1   /* FP32 MFMA in short chains, FP64 accumulation */
#define NOPT_PREC_BF16X3 2 /* bf16 hi/lo split, three MFMA products (about 16-bit products) */
#define NOPT_PREC_BF16X6 3 /* three-piece bf16 split of FP32, six MFMA products */
#define NOPT_PREC_INT8 4   /* int8 digit split with power-of-two row scales, exact int32 MFMA accumulation */

typedef struct nopt_jk_ctx nopt_jk_ctx;

/* Creates a context on HIP device `device` for nbf orbital functions and naux
 * auxiliary functions, with its own stream and the small per-build buffers;
 * B and the per-build tensors are allocated by set_B and the first build.
 * Replaces the allocation part of gen_RI_AA (src/RI.cpp:34-79). */
int nopt_jk_create(int device, long nbf, long naux, nopt_jk_ctx** out);
void nopt_jk_destroy(nopt_jk_ctx* ctx);

/* Which forms of B a context keeps. The default keeps both: FP64 rows (read
 * by the FP64, FP32 and bf16 modes) and the integer forms (read by the
 * integer mode): the packed lower triangle, 5 bytes per triangle element,
 * which the Coulomb pair reads, and, unless the system is large or
 * NOPT_CTX_COMPACT is given, four int8 digit planes of the full rows,
 * 4 bytes per element, which the half-transform reads (faster than
 * unpacking the triangle). The integer forms hold 6.5 bytes per element with
 * the full rows, 2.5 without. A large system keeps one form: FP64 only
 * (8 bytes per element, full rows) or the integer forms only; a build in a
 * mode whose form is not kept fails with NOPT_ESTATE. The full rows are kept
 * when they take at most 24 GiB. NOPT has a single FP64 form of RI_AA
 * (src/RI.cpp:16-32); this choice has no counterpart. */
#define NOPT_CTX_DEFAULT 0
#define NOPT_CTX_FP64_ONLY 1
#define NOPT_CTX_INT8_ONLY 2
#define NOPT_CTX_COMPACT 4 /* or'ed with DEFAULT or INT8_ONLY: the triangle only */
int nopt_jk_create_ex(int device, long nbf, long naux, int flags, nopt_jk_ctx** out);

/* Forms the fitted tensor B = L^-1 (P|mn) on the device in FP64, where
 * V = L L^T is the Cholesky factor of the Coulomb metric, and keeps it
 * resident. eri3 is packed [P][m(m+1)/2+n] (m >= n), metric is naux x naux;
 * eri3 stays in host memory and passes through the device in chunks of rows m
 * of at most 4 GiB (one pass, or two for a digit-planes-only context).
 * NOPT instead keeps raw (mn|P) with V^-1/2 from an eigendecomposition and
 * applies it on every build (src/RI.cpp:57-59, 152-156, src/matr.cpp:372-401);
 * any W with W^T W = V^-1 gives the same J and K. Fails with NOPT_ECHOL when
 * the metric is not positive definite (NOPT drops such eigenvalues with a
 * warning, src/matr.cpp:384-386). */
int nopt_jk_set_B(nopt_jk_ctx* ctx, const double* eri3_packed, const double* metric);

/* Loads an already fitted B (packed [Q][m(m+1)/2+n], FP64) without the metric
 * step. */
int nopt_jk_set_B_fitted(nopt_jk_ctx* ctx, const double* B_packed);

/* Copies the fitted B back to the host (packed layout, FP64). */
int nopt_jk_get_B(nopt_jk_ctx* ctx, double* B_packed);

/* Device bytes held by the context. */
long nopt_jk_device_bytes(const nopt_jk_ctx* ctx);

/* The share of B's integer forms that the integer build reads, after the
 * blocks holding only exact zeros (values below their row's resolution, in
 * every B_Q) are left out: out[0] the 16-column segments and out[1] the
 * 8-column slots of the lower triangle that the Coulomb pair walks, out[2]
 * the 64-column slabs of 32 rows per function that the half-transform
 * multiplies (1.0 where a form is absent). No counterpart in NOPT, whose
 * GEMMs read every element; a diagnostic of the kept forms. */
int nopt_jk_zero_blocks(const nopt_jk_ctx* ctx, double out[3]);

/* Active-space integrals through the fitted B, FP64 (the MO_INTS half of
 * calc_2el_MO_INTS_RI, src/RI.cpp:197-226):
 *   ints[(a2*na + a1)*na*na + b2*na + b1] = sum_Q R_Q[a2][a1] R_Q[b2][b1],
 *   R_Q[a2][a1] = sum_mn C1[a1][m] B_Q,mn C2[a2][n],
 * C1 and C2 na x nbf with orbitals as rows, 1 <= na <= 32. Needs the FP64
 * form of B (NOPT_ESTATE for a context made with NOPT_CTX_INT8_ONLY). Host
 * pointers; synchronous. */
int nopt_jk_active_ints(nopt_jk_ctx* ctx, const double* C1, const double* C2, long na, double* ints);

/* One RI-J/K build, host buffers: J_mn = sum_Q B_Qmn sum_rs B_Qrs D_rs and
 * K_mn = sum_Q sum_{i<nocc} (sum_r B_Qmr C_ir)(sum_s B_Qns C_is).
 * J and K (nbf x nbf) are overwritten. Either output may be NULL to skip it;
 * D may be NULL when J is NULL and C may be NULL when K is NULL.
 * Replaces DM_to_F_transform_RI (src/RI.cpp:115-176). */
int nopt_jk_build(nopt_jk_ctx* ctx, const double* D, const double* C, long nocc, double* J,
                  double* K, int prec);

/* The same build on device buffers, enqueued on `stream` (a hipStream_t; NULL
 * is the default (null) stream, nopt_jk_stream() the context's own); returns
 * after enqueueing. */
int nopt_jk_build_device(nopt_jk_ctx* ctx, const double* dD, const double* dC, long nocc,
                         double* dJ, double* dK, int prec, void* stream);

/* The build of nopt_jk_build_device without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in NOPT. Its work
 * buffers (staged D and C, X and its digit planes, the gamma, J and K
 * partials) live in one static device array of nopt_jk_static_scratch_bytes()
 * bytes (256 MiB) per GPU, held from the library's first use of that GPU
 * until the process ends; a build that needs more fails with NOPT_EINVAL
 * before any device work (e.g. C20H42 in FP64 needs about 0.9 GB: use
 * nopt_jk_build_device). Every call on a GPU shares that GPU's array, so
 * such calls that overlap on one GPU (different streams, different contexts)
 * race; calls on one stream run one after the other. Results equal those of
 * nopt_jk_build_device. */
int nopt_jk_build_device_static(nopt_jk_ctx* ctx, const double* dD, const double* dC, long nocc,
                                double* dJ, double* dK, int prec, void* stream);
long nopt_jk_static_scratch_bytes(void);

/* Several builds against the same B in one pass (NOPT calls the build twice
 * per CASSCF macro-iteration, for the inactive and the total density,
 * src/CAS.cpp:696-707, and a signed factor of an indefinite density needs a
 * positive and a negative orbital factor). Density d uses D[d], the first
 * nocc[d] rows of C[d], and writes J[d] and K[d]; the array pointers J, K, D,
 * C may be NULL and so may individual entries, with the rules of
 * nopt_jk_build per entry. The orbital factors are stacked so that the
 * half-transform reads B once, and gamma and J are formed for all densities
 * in one pass over B each. 1 <= nd <= NOPT_MAX_BATCH. */
#define NOPT_MAX_BATCH 4
int nopt_jk_build_batch(nopt_jk_ctx* ctx, int nd, const double* const* D, const double* const* C,
                        const long* nocc, double* const* J, double* const* K, int prec);
/* The batched build on device buffers, enqueued on `stream` (NULL: the
 * default stream). */
int nopt_jk_build_batch_device(nopt_jk_ctx* ctx, int nd, const double* const* dD,
                               const double* const* dC, const long* nocc, double* const* dJ,
                               double* const* dK, int prec, void* stream);

/* A ragged batch of systems in one call (no counterpart in NOPT, which builds
 * one system at a time): system k builds the nd[k] densities of its own
 * context ctxs[k], items as in nopt_jk_build_batch_device; the item arrays
 * (dD, dC, nocc, dJ, dK) are flattened over the systems in order, system k's
 * items following those of systems 0..k-1. Every argument is checked before
 * anything is enqueued; the systems' builds are then enqueued on `stream` one
 * after the other, so a capture on `stream` records the whole batch as one
 * HIP graph. All contexts are on one device. */
int nopt_jk_build_ragged_device(int nsys, nopt_jk_ctx* const* ctxs, const int* nd, const double* const* dD,
                                const double* const* dC, const long* nocc, double* const* dJ,
                                double* const* dK, int prec, void* stream);

/* A ragged batch as a plan (no counterpart in NOPT; its CASSCF builds one
 * system at a time, src/CAS.cpp:461, 696-707): the integer builds of the
 * systems, items as in nopt_jk_build_ragged_device, are laid out once, and
 * each phase of all of them (staging, half-transform, gamma, J, exchange,
 * outputs) then runs as one launch that walks a work list of (system,
 * density, tile) items, so that small systems fill the device together.
 * create grows the contexts' work buffers, records the work lists in device
 * memory and pins the contexts like a captured graph (no growth or tuning
 * change until destroy); launch enqueues the phase kernels on `stream` (a
 * capture on `stream` records them as one graph: they allocate nothing);
 * the plan reads the inputs and writes the outputs named at create on every
 * launch. Integer mode only; each system's orbitals must fit the fused
 * exchange split (at most 8 segments of 96 columns, NOPT_EINVAL otherwise).
 * Contexts on one device, each at most once. */
typedef struct nopt_ragged nopt_ragged;
int nopt_ragged_create(int nsys, nopt_jk_ctx* const* ctxs, const int* nd, const double* const* dD,
                       const double* const* dC, const long* nocc, double* const* dJ, double* const* dK,
                       nopt_ragged** out);
int nopt_ragged_launch(nopt_ragged* plan, void* stream);
/* Kernel launches per nopt_ragged_launch. */
int nopt_ragged_launches(const nopt_ragged* plan);
void nopt_ragged_destroy(nopt_ragged* plan);

/* Fock assembly and energy on the device from host buffers:
 * F = H + cj*J - ck*K and E = sum D.*F + sum D.*H (E_1el_calc,
 * src/doCI_matr.cpp:38-47, as used at src/SCF.cpp:134; the caller adds
 * V_nuc). F may be NULL; E may be NULL. */
int nopt_jk_fock_energy(nopt_jk_ctx* ctx, const double* H, const double* D, const double* J,
                        const double* K, double cj, double ck, double* F, double* E);

/* Closed-shell energy and its orbital gradient from the occupied factor:
 * D = C^T C (first nocc rows), F = H + 2J - K, E = sum D.*(H+F), and
 * G = dE/dC = 4 C F (nocc x nbf, rows are orbitals). This is the adjoint of
 * E_RHF(C); its occupied-virtual projection is NOPT's orbital gradient
 * g_ia = 4 F_ai (src/converger_2_1.cpp:54-62). G and F (nbf x nbf) may be
 * NULL. */
int nopt_jk_energy_gradient(nopt_jk_ctx* ctx, const double* H, const double* C, long nocc,
                            int prec, double* E, double* G, double* F);
/* The same on device buffers, enqueued on `stream` (NULL: the default
 * stream): E is one device double;
 * G (nocc x nbf) and F (nbf x nbf) may be NULL. The context's J, K and D
 * buffers are used as scratch. */
int nopt_jk_energy_gradient_device(nopt_jk_ctx* ctx, const double* dH, const double* dC, long nocc,
                                   int prec, double* dE, double* dG, double* dF, void* stream);

/* Tile and split tuning. The build divides the auxiliary index into Q-splits
 * (partial sums over Q ranges, added in FP64 at the end): params[6] =
 * {FP64/FP32/bf16 exchange splits, FP64 J splits (0: the fewest the kernel's
 * gamma cache allows, at least 16), integer exchange splits, integer gamma
 * splits, integer J splits, and where the integer Coulomb
 * pair starts on its side stream: 0 after the half-transform, 1 gamma beside
 * the half-transform and J after it, 2 both beside it}. create picks them from the
 * device's CU count; nopt_jk_autotune times `reps` device builds of (dD, dC)
 * in mode `prec` on `stream` for each candidate of each split count in turn
 * (two rounds of coordinate descent), keeps the fastest and returns its mean
 * milliseconds per build in *best_ms (may be NULL). Results differ from the untuned build
 * only in the order of the FP64 partial sums. No counterpart in NOPT, whose
 * DGEMMs split rows evenly over threads (include/blas_link.h:54-116). */
int nopt_jk_get_tuning(const nopt_jk_ctx* ctx, int* params);
int nopt_jk_set_tuning(nopt_jk_ctx* ctx, const int* params);
int nopt_jk_autotune(nopt_jk_ctx* ctx, const double* dD, const double* dC, long nocc, int prec, int reps,
                     double* best_ms, void* stream);

/* Per-phase timing with HIP events on the build's stream (no counterpart in
 * NOPT, whose RHF loop disables its step timers, src/SCF.cpp:124). When
 * enabled, each build (or batch) records events around its six phases:
 * staging of D and C, half-transform (with the fused gamma partials of a
 * single FP64/FP32/bf16 build), gamma, J partials, exchange (K products and
 * K output) and the J output. read_timing returns the summed milliseconds per
 * phase (ms[6]) and the number of builds recorded since the last read, then
 * resets. */
int nopt_jk_set_timing(nopt_jk_ctx* ctx, int enable);
int nopt_jk_read_timing(nopt_jk_ctx* ctx, double* ms, long* builds);
/* The context's own stream (hipStream_t). */
void* nopt_jk_stream(nopt_jk_ctx* ctx);

/* HIP graph of device builds (no counterpart in NOPT): between begin and end,
 * nopt_jk_build_device / nopt_jk_build_batch_device calls on `stream` (of
 * one or several contexts on the device, e.g. a ragged batch of molecules)
 * are recorded instead of run; launch replays them all with one call. The
 * inputs and outputs are the device buffers named at capture time (update
 * their contents between launches). Run one build of each shape before
 * capturing: buffers cannot grow during a capture (NOPT_ESTATE).
 * A graph holds the addresses of the work buffers of every context it
 * recorded, so from capture_end until nopt_graph_destroy those contexts are
 * pinned: a build that would grow a buffer (more occupied orbitals, a larger
 * batch, another mode), nopt_jk_set_tuning and nopt_jk_autotune fail with
 * NOPT_ESTATE, and nopt_jk_destroy only marks the context, which is freed
 * when its last graph is destroyed. A launch waits for the earlier work of
 * the recorded contexts and later calls on them wait for the launch. */
typedef struct nopt_graph nopt_graph;
int nopt_graph_capture_begin(int device, void* stream);
int nopt_graph_capture_end(int device, void* stream, nopt_graph** out);
int nopt_graph_launch(nopt_graph* graph, void* stream);
void nopt_graph_destroy(nopt_graph* graph);

/* Device-memory plumbing for callers without a HIP runtime binding.
 * kind: 0 host->device, 1 device->host, 2 device->device. */
void* nopt_dev_malloc(int device, size_t bytes);
void nopt_dev_free(void* ptr);
int nopt_dev_memcpy(void* dst, const void* src, size_t bytes, int kind);
int nopt_dev_synchronize(int device);

const char* nopt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
