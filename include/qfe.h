/*
 * qfe.h -- C ABI of the MI355X Fock-build engine (ERIs + Schwarz screening +
 * J/K digestion) for qdk-chemistry's SCF solver.
 *
 * It is the drop-in for the reference's integral-direct J/K engine
 * (cpp/src/qdk/chemistry/algorithms/microsoft/scf/src/eri/LIBINT2_DIRECT/
 * libint2_direct.cpp) behind the abstract qdk::chemistry::scf::ERI interface
 * (scf/include/qdk/chemistry/scf/core/eri.h:23-176), in the slot the reference
 * reserves for a GPU engine (ERIMethod::Rys, core/enums.h:22-35, created in
 * scf/src/eri/eri.cpp:36-68).  Plain pointers and sizes only.
 *
 * Conventions (match the reference):
 *   - AO order is the order of the shells passed in qfe_basis (the reference
 *     sorts shells by L, then descending leading exponent), functions within
 *     a shell: cartesian CCA order for l < 2, solid harmonics m = -l..l for
 *     l >= 2 (pure bases).
 *   - Contraction coefficients are used as given (the reference builds its
 *     libint2 shells with embed_normalization=false, libint2_util.cpp:15-28).
 *   - Matrices are row-major double, ndm x nbf x nbf.
 *   - J and K come back symmetrised; K includes the (alpha + beta) scale.
 *   - Gradient outputs are 3*natoms, xyz-major: index = atom + xyz*natoms
 *     (libint2_direct.cpp:824).
 * Every function returns 0 on success and a non-zero status otherwise; the
 * message is available from qfe_last_error() (thread-local).
 */
#ifndef QFE_H
#define QFE_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define QFE_OK 0
#define QFE_ERR_INVALID 1     /* invalid argument (std::invalid_argument in the reference) */
#define QFE_ERR_RUNTIME 2     /* runtime failure (std::runtime_error), e.g. NaN density */
#define QFE_ERR_DEVICE 3      /* HIP runtime error */
#define QFE_ERR_UNSUPPORTED 4 /* feature outside this engine (e.g. range-separated exchange) */

typedef struct qfe_handle qfe_handle;

/* Basis set in AO order.  Replaces qdk::chemistry::scf::BasisSet
 * (scf/include/qdk/chemistry/scf/core/basis_set.h:30-58) at the boundary. */
typedef struct {
  int nshells;
  const int* l;            /* [nshells] angular momentum */
  const int* atom;         /* [nshells] atom index */
  const double* centers;   /* [nshells][3] bohr */
  const int* prim_offset;  /* [nshells + 1] offsets into exps/coefs */
  const double* exps;      /* primitive exponents */
  const double* coefs;     /* contraction coefficients (normalisation included) */
  int pure;                /* 1: solid harmonics for l >= 2 */
  int natoms;
  const double* charges;   /* [natoms] nuclear charges (one-electron integrals) */
  const double* coords;    /* [natoms][3] bohr */
} qfe_basis;

/* Engine settings.  Field names follow ERIConfig (core/scf.h:65-94). */
typedef struct {
  double eri_threshold;        /* Schwarz x density screening threshold (1e-12 plugin default) */
  double shell_pair_threshold; /* overlap threshold for significant pairs (1e-12) */
  int device_id;               /* HIP device ordinal */
  int mix_precision;           /* 1: FP32 for quartets whose Schwarz bound < mix_threshold */
  double mix_threshold;        /* 1e-6 */
  int verbose;
  int profile;                 /* 1: time every class kernel group with HIP events (its class kernels then
                                  run one after another; else they spread over QFE_STREAMS streams) */
  int use_atomics;             /* 0 (the reference's default, core/scf.h:81-82): bitwise-repeatable
                                  J/K (order-independent fixed-point accumulation); 1: FP64 atomics */
} qfe_config;

/* Per-build statistics (the last build). */
typedef struct {
  long long quartets;          /* shell quartets that passed screening */
  long long quartets_fp32;     /* of which evaluated in FP32 */
  long long prim_quartets;     /* primitive quartets of the evaluated shell quartets */
  double ms_screen;            /* task-list construction */
  double ms_eri;               /* ERI + digestion kernels */
  double ms_total;             /* whole build, device time */
  long long significant_pairs;
} qfe_stats;

/* Create the engine: significant pairs, primitive-pair tables, Schwarz.
 * Replaces libint2::direct::ERI::ERI (libint2_direct.cpp:204-226). */
int qfe_create(const qfe_basis* basis, const qfe_config* cfg, qfe_handle** out);

/* J and K for ndm host densities (1: RHF total density; 2: the spin pair
 * [P_alpha; P_beta], both screened with the norms of P_alpha as the
 * reference's loop does, libint2_direct.cpp:271).  Replaces ERI::build_JK
 * (eri.h:79-80; libint2_direct.cpp:256-626).  J or K may be NULL: that
 * matrix is not built (ERIMultiplexer's split engines,
 * eri_multiplexer.cpp:66-67).  omega != 0 is unsupported. */
int qfe_build_jk(qfe_handle* h, const double* P, double* J, double* K, double alpha, double beta,
                 double omega, int ndm);

/* Same on device-resident buffers (HIP device pointers, nbf x nbf each);
 * stream may be NULL for the engine's own stream.  J or K may be NULL. */
int qfe_build_jk_device(qfe_handle* h, const double* dP, double* dJ, double* dK, double alpha,
                        double beta, double omega, void* stream);

/* qfe_build_jk_device without any device allocation in the call.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-build task list, J_cd share slots and survivor masks
 * live in a fixed device array of qfe_noalloc_scratch_bytes() (256 MiB) per
 * GPU, held until the process ends; a build that needs more returns
 * QFE_ERR_INVALID (nothing is computed).  Every call on one GPU shares that
 * array: calls that overlap on one GPU race, and the array keeps the last
 * call's contents (each build clears what it needs).  Arguments and results
 * are those of qfe_build_jk_device. */
int qfe_build_jk_device_noalloc(qfe_handle* h, const double* dP, double* dJ, double* dK, double alpha,
                                double beta, double omega, void* stream);
unsigned long long qfe_noalloc_scratch_bytes(void);

/* Engine-owned HBM buffers of nbf x nbf doubles for callers that keep P, J
 * and K resident on the device (e.g. to time qfe_build_jk_device without
 * PCIe traffic), and a synchronous copy on the engine's stream
 * (kind 0: host -> device, 1: device -> host, 2: device -> device through a
 * plain copy kernel; kind 2 needs 8-byte aligned pointers and a multiple of 8
 * bytes).  No reference counterpart: the reference's build_JK takes host
 * pointers only. */
int qfe_device_buffers(qfe_handle* h, double** dP, double** dJ, double** dK);
int qfe_copy(qfe_handle* h, void* dst, const void* src, unsigned long long bytes, int kind);

/* ERI nuclear-gradient contributions dJ, dK (3*natoms, xyz-major) for ndm
 * densities: 1 = the RHF total density; 2 = the spin pair [P_alpha; P_beta],
 * whose Coulomb term takes P_alpha + P_beta and whose exchange term sums the
 * spins with scale -1/2 (libint2_direct.cpp:683-692,866-886).  Quartets are
 * screened with the norms of the first density block.  A NULL dJ or dK is
 * not computed; the outputs are overwritten.  Replaces ERI::get_gradients
 * (eri.h:100-101; libint2_direct.cpp:646-890). */
int qfe_gradients(qfe_handle* h, const double* P, double* dJ, double* dK, double alpha,
                  double beta, double omega, int ndm);

/* One-electron matrices (host): op 0 = overlap, 1 = kinetic, 2 = nuclear
 * attraction.  Harness support for the SCF driver (scf/src/util/int1e.cpp). */
int qfe_one_body(const qfe_basis* basis, int op, double* out);

/* Gradient of sum_mn W_mn d(O)_mn / dR for op 0/1/2 (3*natoms, atom-major). */
int qfe_one_body_grad(const qfe_basis* basis, int op, const double* W, double* grad);

/* The same two on GPU device_id (host inputs and outputs; shells up to f).
 * Replace OneBodyIntegral::overlap_integral / kinetic_integral /
 * nuclear_integral (scf/include/qdk/chemistry/scf/util/int1e.h; Libint2
 * evaluation in scf/src/util/int1e.cpp:473-977) and their *_deriv(dm, res)
 * contractions (the same file).  The nuclear charges are those of the basis
 * struct: Z minus the ECP core electrons, as the reference's atomic_charges
 * (core/basis_set.cpp:85-87). */
int qfe_one_body_gpu(const qfe_basis* basis, int op, double* out, int device_id);
int qfe_one_body_grad_gpu(const qfe_basis* basis, int op, const double* W, double* grad, int device_id);

/* Eigenvalues w (ascending) and eigenvectors V (row-major, eigenvector j in
 * column j) of the symmetric n x n matrix A, on GPU device_id with
 * rocSOLVER dsyevd.  Stands in for the SCF's diagonalisation of the
 * orthogonalised Fock matrix (the reference's optional device path,
 * scf_algorithm.cpp:105-152, scf_impl.cpp:431-449, and its LAPACK one). */
int qfe_eigh_gpu(int n, const double* A, double* w, double* V, int device_id);

/* Density-fitted Coulomb matrix on GPU device_id.  Replaces the J of the
 * reference's INCORE DF engine (ERI_DF, scf/src/eri/INCORE/
 * incore_impl_df.cpp:53-200; eri_df_base.cpp): create computes the
 * three-centre integrals (mu nu|P) over the auxiliary basis `aux` and the
 * Cholesky factor of the metric (P|Q) once (orbital shells up to d,
 * auxiliary shells up to g); build_j returns J = B V^-1 B^T vec(P) for a
 * host density (nbf x nbf).  K is not fitted: the multiplexer pairs this J
 * with the exchange of the Rys engine (eri_multiplexer.cpp:34-71). */
typedef struct qfe_dfj qfe_dfj;
int qfe_dfj_create(const qfe_basis* basis, const qfe_basis* aux, int device_id, qfe_dfj** out);
int qfe_dfj_build_j(qfe_dfj* h, const double* P, double* J);
int qfe_dfj_naux(const qfe_dfj* h);
void qfe_dfj_destroy(qfe_dfj* h);

/* First quarter of the MO transform, ERI::quarter_trans(nt, C, out)
 * (eri.h:116; libint2_direct.cpp:913-1141, incore_impl.cpp:242-297):
 * out[i][j][k][p] = sum_l (ij|kl) C[l][p], C row-major nbf x nt, out
 * nbf^3 x nt with p fastest (host memory).  The dense AO tensor is computed
 * in HBM (shells up to d) and contracted with one rocBLAS GEMM; nbf^4
 * doubles must fit in free device memory (QFE_ERR_UNSUPPORTED otherwise).
 * Requires 0 < nt <= nbf. */
int qfe_quarter_trans(qfe_handle* h, unsigned long long nt, const double* C, double* out);

/* Four-index MO transform, MOERI::compute(nb, nt, Ci, Cj, Ck, Cl, out)
 * (moeri.cpp:28-153): out[w][x][y][z] = sum (ab|cd) Cl[a][w] Ck[b][x]
 * Cj[c][y] Ci[d][z], each C row-major nb x nt, out nt^4 (host).  nb must be
 * the basis size; nt > nb returns QFE_ERR_RUNTIME with the reference's
 * message.  Four rocBLAS GEMMs over the dense tensor in HBM. */
int qfe_mo_eri(qfe_handle* h, unsigned long long nb, unsigned long long nt, const double* Ci, const double* Cj,
               const double* Ck, const double* Cl, double* out);

/* Device-resident RHF iterations: the dense linear algebra of
 * SCFImpl::iterate_ (scf_impl.cpp:555-621) and the DIIS of
 * scf_algorithm.cpp:401-435 kept in the engine's HBM, replacing the host
 * GEMMs and the per-iteration P/J/K copies (the reference's optional device
 * path is the cuSOLVER eigensolve, scf_algorithm.cpp:105-152).
 *   create: H, S (nbf x nbf), X (nbf x norb, row-major, the orthogonaliser),
 *           nocc doubly occupied orbitals, E_nuc, DIIS depth.
 *   set_density: the guess P (host).
 *   step: J/K of P (full_build != 0: F = H + J - 1/2 K) or of P - P_last
 *         (F += J - 1/2 K); returns E = E_nuc + 1/2 Tr(PH) + 1/2 Tr(PF), the
 *         orbital gradient max|FPS - SPF| / nbf and ||P - P_prev||_F / nbf.
 *   update: DIIS extrapolation over the stored (F, FPS - SPF), X^T F X,
 *           rocSOLVER dsyevd, P = 2 C_occ C_occ^T.
 *   finish: a full rebuild, F, E and the eigenvectors of that F.
 *   get: what 0 = P, 1 = F (nbf x nbf), 2 = C (nbf x norb, row-major),
 *        3 = orbital energies (norb).
 * The context uses the engine's stream and device buffers: do not run other
 * builds on the same handle between step and update. */
typedef struct qfe_scf qfe_scf;
int qfe_scf_create(qfe_handle* h, const double* H, const double* S, const double* X, int norb, int nocc,
                   double enuc, int diis_size, qfe_scf** out);
int qfe_scf_set_density(qfe_scf* s, const double* P);
int qfe_scf_step(qfe_scf* s, int full_build, double* energy, double* orbital_gradient, double* density_rms);
int qfe_scf_update(qfe_scf* s);
int qfe_scf_finish(qfe_scf* s, double* energy);
int qfe_scf_get(qfe_scf* s, int what, double* out);
void qfe_scf_destroy(qfe_scf* s);

/* Switch the precision policy of subsequent builds: mix != 0 evaluates
 * quartets with Q_ab Q_cd < threshold in FP32 (ERIConfig::is_mix_precision,
 * core/scf.h:71-93, which the reference declares but no shipped engine reads);
 * mix == 0 is all-FP64.  The SCF driver uses it for FP64 final iterations. */
int qfe_set_precision(qfe_handle* h, int mix, double threshold);

int qfe_get_stats(const qfe_handle* h, qfe_stats* out);

/* Per-class census of the last build for precision 0 (FP64 tasks) or 1
 * (FP32 tasks of the mixed-precision path): for class pair i < min(cap, n):
 * cls[4i..4i+3] = (la, lb, lc, ld) as evaluated, quartets[i] and
 * prim_quartets[i] the screened shell quartets and their primitive quartets
 * over significant primitive pairs, exec_prim_quartets[i] (may be NULL) the
 * primitive quartets the kernels evaluated after primitive screening
 * (counted on this call by a counting kernel over the last J/K build's task
 * list, synchronously on the engine's stream; 0 after a gradient call), and
 * ms[i] the kernel time of that class (the last two only with
 * qfe_config.profile, else 0; a profiled engine launches its class kernels
 * one after another).  Returns n (number of class pairs), or -1 for a null
 * handle or a failed count.  No reference counterpart: instrumentation. */
int qfe_class_stats(const qfe_handle* h, int precision, int cap, int* cls, double* quartets,
                    double* prim_quartets, double* exec_prim_quartets, double* ms);
/* Chunk geometry of the J/K kernels of class (la lb | lc ld) as compiled:
 * returns the number of chunks n and, for i < min(n, cap), out[4i..4i+3] =
 * the chunk's bra functions [Q0, Q1) and ket cartesians [K0, K1); -1 for a
 * class that is not compiled.  No reference counterpart: instrumentation, the
 * input of the FLOP model behind the benchmark's roofline. */
int qfe_class_chunks(int la, int lb, int lc, int ld, int cap, int* out);
int qfe_nbf(const qfe_handle* h);
/* Highest shell angular momentum compiled into the ERI kernels of this build
 * (qfe_create rejects bases with higher l).  No reference counterpart. */
int qfe_max_l(void);
void qfe_destroy(qfe_handle* h);
const char* qfe_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QFE_H */
