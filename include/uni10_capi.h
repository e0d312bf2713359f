/*
 * uni10_capi.h - C facade of the Uni10 C++ API (libuni10_amd.so).
 *
 * The binding surface a foreign-language wrapper uses, in the role of the
 * reference's SWIG module (uni10/Swig/pyUni10.i): opaque handles for
 * UniTensor and Network, host arrays in and out, status codes instead of
 * exceptions (u10_last_error() holds the message).  Every call runs the C++
 * API, which runs the HIP kernels behind tn_hip.h.
 *
 * Qnums are passed as int triples {U1, prt, prtF}.
 */
#ifndef UNI10_CAPI_H
#define UNI10_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct u10_tensor_s* u10_tensor;
typedef struct u10_network_s* u10_network;
typedef struct u10_itebd_s* u10_itebd;
typedef struct u10_batch_tebd_s* u10_batch_tebd;

const char* u10_last_error(void);
int u10_set_gemm_mode(int mode);          /* GemmMode of uni10.hpp (thread-local) */
int u10_set_stream(void* stream);         /* thread-local current stream          */
int u10_set_autotune(int on);             /* tile autotuning of new GEMM plans    */

/* UniTensor(bonds, name): bond b has type types[b] (1 = BD_IN, -1 = BD_OUT)
 * and nruns[b] runs; run r of all bonds, in order, has charge
 * run_qnums[3*r..3*r+2] and degeneracy run_degs[r]. */
int u10_tensor_create(int nbonds, const int* types, const int* nruns, const int* run_qnums,
                      const int* run_degs, const char* name, u10_tensor* out);
/* A complex tensor (UniTensor(CTYPE, bonds), UniTensor.h:798): elements
 * exchanged as interleaved (re, im) doubles by set/get_elem, set/get_raw and
 * put/get_block; u10_tensor_type gives 1 (real) or 2 (complex). */
int u10_tensor_create_c(int nbonds, const int* types, const int* nruns, const int* run_qnums, const int* run_degs,
                        const char* name, u10_tensor* out);
int u10_tensor_type(u10_tensor t, int* type);
/* Bond::combine (Bond.cpp:163-196; use_free_fn = 0: bonds[0].combine(bonds[1])
 * .combine(bonds[2])...) or the free combine(tp, bonds) (Bond.cpp:198-252;
 * use_free_fn = 1).  Host-only (no device needed).  Outputs: type, dim, the
 * run-length charges (up to cap runs, (U1, prt, prtF) triples) and
 * degeneracy() in Qnum order. */
int u10_bond_combine(int nbonds, const int* types, const int* nruns, const int* run_qnums, const int* run_degs,
                     int use_free_fn, int tp, int* out_type, int* out_dim, int* out_nruns, int* out_qnums,
                     int* out_degs, int* out_ndeg, int* deg_qnums, int* deg_counts, int cap);
int u10_tensor_scalar(double value, u10_tensor* out);
int u10_tensor_load(const char* fname, u10_tensor* out);        /* UniTensor(fname) */
int u10_tensor_save(u10_tensor t, const char* fname);
int u10_tensor_copy(u10_tensor t, u10_tensor* out);
int u10_tensor_free(u10_tensor t);
int u10_tensor_info(u10_tensor t, int* nbonds, int* in_bonds, int64_t* elem_num, int* nblocks);
int u10_tensor_labels(u10_tensor t, int* labels);
int u10_tensor_set_label(u10_tensor t, int n, const int* labels);
int u10_tensor_bond(u10_tensor t, int b, int* type, int* dim, int* nruns);
int u10_tensor_bond_runs(u10_tensor t, int b, int* run_qnums, int* run_degs);
int u10_tensor_block(u10_tensor t, int i, int* qnum, int64_t* rows, int64_t* cols, int64_t* offset);
int u10_tensor_set_elem(u10_tensor t, const double* host);     /* block layout */
int u10_tensor_get_elem(u10_tensor t, double* host);
int u10_tensor_set_raw(u10_tensor t, const double* raw);       /* setRawElem */
int u10_tensor_raw_shape(u10_tensor t, int64_t* rows, int64_t* cols);
int u10_tensor_get_raw(u10_tensor t, double* raw);             /* getRawElem */
int u10_tensor_at(u10_tensor t, const int64_t* idxs, double* out);
int u10_tensor_put_block(u10_tensor t, const int* qnum, int64_t rows, int64_t cols, const double* host);
int u10_tensor_get_block(u10_tensor t, const int* qnum, double* host);
int u10_tensor_permute(u10_tensor t, int n, const int* labels, int in_bond_num);
int u10_tensor_transpose(u10_tensor t);
int u10_tensor_scale(u10_tensor t, double a);
/* randomize: elements (block layout order) = SplitMix64 uniform[-1, 1) of
 * (seed, 1, 2, ...), generated on the device (elemRand, uni10_tools.h:101) */
int u10_tensor_randomize(u10_tensor t, uint64_t seed);
int u10_tensor_norm(u10_tensor t, double* out);
int u10_tensor_device_ptr(u10_tensor t, void** dptr);          /* elements in HBM */
int u10_contract(u10_tensor a, u10_tensor b, int fast, u10_tensor* out);
int u10_otimes(u10_tensor a, u10_tensor b, u10_tensor* out);
/* operator*(UniTensor, UniTensor) (UniTensor.cpp:224-234): contract of copies. */
int u10_tensor_mul(u10_tensor a, u10_tensor b, u10_tensor* out);
/* exSwap (UniTensor.cpp:1237-1268): *n gate pairs (b1, b2), the first `cap`
 * written to pairs[2*i], pairs[2*i+1]. */
int u10_tensor_ex_swap(u10_tensor a, u10_tensor b, int* pairs, int cap, int* n);
/* addGate (UniTensorReal.cpp:868-921): n gates as (b1, b2) pairs. */
int u10_tensor_add_gate(u10_tensor t, int n, const int* pairs);

int u10_network_create(const char* fname, u10_network* out);
int u10_network_put(u10_network net, const char* name, u10_tensor t, int transposed);
int u10_network_launch(u10_network net, const char* name, u10_tensor* out);
int u10_network_tree(u10_network net, char* buf, int len);
/* operator<<(ostream, Network) (Network.cpp:856-907): tensor list, then the
 * contraction tree with element counts, in the reference's text format. */
int u10_network_print(u10_network net, char* buf, int len);
/* Network::setGraph: launches replay one captured HIP graph per leaf state
 * (no counterpart in the reference). */
int u10_network_set_graph(u10_network net, int on);
int u10_network_free(u10_network net);

/* RDotR (BlockTools.cpp:38-76), the Matrix product of the C++ API, on host
 * arrays: a diagonal operand is given by its min(rows, cols) diagonal
 * entries; the result (rows x cols, diagonal when both operands are) is
 * written to c (at most cap doubles). */
int u10_rdotr(int64_t ar, int64_t ac, int adiag, const double* a, int64_t br, int64_t bc, int bdiag,
              const double* b, int64_t* cr, int64_t* cc, int* cdiag, double* c, int64_t cap);

/* Matrix::svd / eigh / exph on host arrays (computed on the device). */
int u10_svd(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT);
/* Complex SVD and Hermitian eigen solver through Matrix(CTYPE): A, U, VT,
 * vec interleaved (re, im); S and eig real. */
int u10_svd_c128(const double* A, int64_t M, int64_t N, double* U, double* S, double* VT);
int u10_eigh_c128(const double* A, int64_t N, double* eig, double* vec);
int u10_eigh(const double* A, int64_t N, double* eig, double* vec_rows);
int u10_exph(double a, const double* H, int64_t N, double* out);
/* lanczosEigh (Block.cpp:474-506) on host arrays: psi (N) is the start
 * vector on entry and the ground state on return; *iters = steps taken. */
int u10_lanczos_eigh(const double* A, int64_t N, double* psi, uint64_t max_iter, double err_tol, double* E0,
                     uint64_t* iters);

/* iTEBD of the spin-1/2 XXZ chain through Networks (the example driver).
 * netdir (here and in u10_batch_tebd_create*): the directory holding
 * itebd.net, updateA.net, MPS.net and measure.net; NULL or "" = the nets/
 * directory installed beside libuni10_amd.so. */
int u10_itebd_create(int chi, int d, double delta, double tau, uint64_t seed, const char* netdir,
                     u10_itebd* out);
int u10_itebd_step(u10_itebd t, int nsteps);
/* replay the driver's update networks as captured graphs (Network::setGraph) */
int u10_itebd_set_graph(u10_itebd t, int on);
int u10_itebd_energy(u10_itebd t, double* out);
int u10_itebd_lambda(u10_itebd t, int which, double* out, int* n);
int u10_itebd_trunc_err(u10_itebd t, double* out);
int u10_itebd_free(u10_itebd t);

/* Batched iTEBD: n independent XXZ chains (bond dimension chi[i] <= 2048,
 * anisotropy delta[i], seed[i]) advanced together, each update stage one
 * launch for all chains; gemm_mode is a GemmMode of uni10.hpp (SVD and
 * truncation stay FP64).  set_graph(1) replays steps as a captured graph.
 * Per chain, the update of the example driver (iTEBD.cpp:148-173). */
int u10_batch_tebd_create(int n, const int* chi, const double* delta, const uint64_t* seed, double tau,
                          int gemm_mode, const char* netdir, u10_batch_tebd* out);
/* As create, with per-chain physical dimension d[i] (2: spin-1/2 XXZ, 3:
 * spin-1 XXZ; NULL = all 2) and symmetry sym[i] (0: dense, 1: U(1) with
 * conserved Sz, block-sparse tensors, per-charge-block SVD and fixed sector
 * dimensions; NULL = all dense).  No counterpart in the reference, whose
 * iTEBD example is dense spin-1/2; a U(1) chain is that update run per
 * Qnum block (Block::svd per block, BlockReal.cpp:226-250). */
int u10_batch_tebd_create_ex(int n, const int* chi, const double* delta, const uint64_t* seed, const int* d,
                             const int* sym, double tau, int gemm_mode, const char* netdir, u10_batch_tebd* out);
/* Charge sectors of bond `bond` (0: lambda_A, 1: lambda_B) of a U(1) chain:
 * sets *count to the number of sectors (0 for a dense chain).  With charges
 * or dims non-NULL, *count must hold their capacity on input, and the
 * sectors' charges (2 Sz, ascending) and dimensions are written. */
int u10_batch_tebd_sectors(u10_batch_tebd b, int chain, int bond, int* charges, int* dims, int* count);
/* The sector rule itself (host only, no device work): the charges and
 * dimensions of bond `bond` of a U(1) chain of bond dimension chi and
 * physical dimension d, with the same count/capacity convention. */
int u10_u1_bond_sectors(int chi, int d, int bond, int* charges, int* dims, int* count);
int u10_batch_tebd_step(u10_batch_tebd b, int nsteps);
/* step without waiting: each engine has its own stream, so the steps of
 * several engines overlap on the device; wait finishes the engine's work and
 * checks that its SVDs converged */
int u10_batch_tebd_launch(u10_batch_tebd b, int nsteps);
int u10_batch_tebd_wait(u10_batch_tebd b);
/* *ready = 1 once the launched steps have finished, 0 while they run (a
 * poll that never blocks; wait still checks the SVDs) */
int u10_batch_tebd_ready(u10_batch_tebd b, int* ready);
int u10_batch_tebd_set_graph(u10_batch_tebd b, int on);
/* The engine's work runs on `stream` (NULL: its own) from now on; waits for
 * the engine to be idle first. */
int u10_batch_tebd_set_stream(u10_batch_tebd b, void* stream);
/* chain on/off: an inactive chain is skipped on the device by every launch
 * (its tensors, spectra and warm-start data are kept) */
int u10_batch_tebd_set_active(u10_batch_tebd b, int chain, int on);
/* restart one chain (slot reuse): the seeded start of (delta, seed) */
int u10_batch_tebd_load(u10_batch_tebd b, int chain, double delta, uint64_t seed);
int u10_batch_tebd_energy(u10_batch_tebd b, int chain, double* out);
int u10_batch_tebd_lambda(u10_batch_tebd b, int chain, int which, double* out, int* n);
int u10_batch_tebd_trunc_err(u10_batch_tebd b, int chain, double* out);
int u10_batch_tebd_info(u10_batch_tebd b, double* gemm_flops_per_step, int* last_sweeps);
int u10_batch_tebd_free(u10_batch_tebd b);

#ifdef __cplusplus
}
#endif
#endif /* UNI10_CAPI_H */
