/*
 * CPU oracle for the Uni10 index-remap hot path.  TEST INFRASTRUCTURE ONLY:
 * nothing in the product links or calls this file; tests and the bench's
 * cpu_baseline leg load it through ctypes as the checker.
 *
 * Each routine restates one reference loop, element for element:
 *
 *   ora_permute_dense  - non-symmetric permute, host branch
 *                        (uni10/src/uni10/tensor-network/lib/UniTensorReal.cpp:601-646):
 *                        walk the source in storage order with an odometer over the
 *                        source bond indices and scatter to the destination offset
 *                        kept incrementally in `cnt_ot` (transAcc strides).
 *   ora_reshape_elem   - the device remap formula of _reshapeElem
 *                        (uni10/src/uni10/tools/lib/uni10_tools_gpu.cu:319-330) with the
 *                        offset table built at UniTensorReal.cpp:586-595.
 *   ora_permute_moves  - symmetric permute, per-Qidx sub-block copy
 *                        (UniTensorReal.cpp:739-755): for every sub-block, walk the
 *                        source sub-block row-major, keep the destination linear
 *                        index with the sBot_acc odometer, and place it at
 *                        (cnt_ot / sBot_cDim, cnt_ot % sBot_cDim) in the output block.
 *   ora_gemm_rowmajor  - matrixMul (uni10_lapack_cpu.cpp:41-44): row-major
 *                        C(MxN) = A(MxK) B(KxN), plain triple loop in double.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

/* rsp_outin[j] = source bond placed at output position j. */
int ora_permute_dense(const double* src, double* dst, int rank,
                      const int64_t* dims, const int* rsp_outin) {
  if (rank <= 0 || rank > 32) return -1;
  int64_t new_dims[32], new_acc[32], trans_acc[32], idx[32];
  int64_t total = 1;
  for (int b = 0; b < rank; ++b) {
    new_dims[b] = dims[rsp_outin[b]];
    total *= dims[b];
  }
  new_acc[rank - 1] = 1;
  for (int b = rank - 1; b > 0; --b) new_acc[b - 1] = new_acc[b] * new_dims[b];
  for (int b = 0; b < rank; ++b) {
    trans_acc[rsp_outin[b]] = new_acc[b];
    idx[b] = 0;
  }
  int64_t cnt_ot = 0;
  for (int64_t i = 0; i < total; ++i) {
    dst[cnt_ot] = src[i];
    for (int bend = rank - 1; bend >= 0; --bend) {
      idx[bend]++;
      if (idx[bend] < dims[bend]) {
        cnt_ot += trans_acc[bend];
        break;
      }
      cnt_ot -= trans_acc[bend] * (idx[bend] - 1);
      idx[bend] = 0;
    }
  }
  return 0;
}

int ora_reshape_elem(const double* src, double* dst, int rank,
                     const int64_t* dims, const int* rsp_outin) {
  if (rank <= 0 || rank > 32) return -1;
  int64_t off[64], new_acc[32];
  int64_t total = 1;
  for (int b = 0; b < rank; ++b) total *= dims[b];
  new_acc[rank - 1] = 1;
  off[rank - 1] = 1;
  for (int b = rank - 1; b > 0; --b) {
    new_acc[b - 1] = new_acc[b] * dims[rsp_outin[b]];
    off[b - 1] = off[b] * dims[b];
  }
  for (int b = 0; b < rank; ++b) off[rank + rsp_outin[b]] = new_acc[b];
  for (int64_t old_idx = 0; old_idx < total; ++old_idx) {
    int64_t rem = old_idx, new_idx = 0;
    for (int i = 0; i < rank; ++i) {
      new_idx += (rem / off[i]) * off[rank + i];
      rem = rem % off[i];
    }
    dst[new_idx] = src[old_idx];
  }
  return 0;
}

/*
 * One sub-block move per entry.  Integer metadata per move m (all int64):
 *   meta[m*6 + 0] src offset of the sub-block corner in `src`
 *   meta[m*6 + 1] src block column count (Bin_cDim, the row pitch)
 *   meta[m*6 + 2] src sub-block column count (sBin_cDim)
 *   meta[m*6 + 3] dst offset of the sub-block corner in `dst`
 *   meta[m*6 + 4] dst block column count (Bot_cDim)
 *   meta[m*6 + 5] dst sub-block column count (sBot_cDim)
 * sub_dims[m*rank + b]  per-source-bond sub-block extent (sBin_sBdims)
 * bot_acc[m*rank + b]   destination linear stride of source bond b (sBot_acc)
 * sign[m]               +1 or -1 (fermionic swap sign)
 */
int ora_permute_moves(const double* src, double* dst, int rank, int64_t nmoves,
                      const int64_t* meta, const int64_t* sub_dims,
                      const int64_t* bot_acc, const double* sign) {
  if (rank <= 0 || rank > 32) return -1;
  int64_t sidx[32];
  for (int64_t m = 0; m < nmoves; ++m) {
    const int64_t* mm = meta + m * 6;
    const int64_t* sd = sub_dims + m * rank;
    const int64_t* acc = bot_acc + m * rank;
    const double* ein = src + mm[0];
    double* eot = dst + mm[3];
    int64_t bin_c = mm[1], sbin_c = mm[2], bot_c = mm[4], sbot_c = mm[5];
    int64_t total = 1;
    for (int b = 0; b < rank; ++b) {
      total *= sd[b];
      sidx[b] = 0;
    }
    int64_t sbin_r_dim = sbin_c ? total / sbin_c : 0;
    int64_t cnt_ot = 0;
    for (int64_t r = 0; r < sbin_r_dim; ++r)
      for (int64_t c = 0; c < sbin_c; ++c) {
        int64_t orow = cnt_ot / sbot_c, ocol = cnt_ot % sbot_c;
        eot[orow * bot_c + ocol] = sign[m] * ein[r * bin_c + c];
        for (int bend = rank - 1; bend >= 0; --bend) {
          sidx[bend]++;
          if (sidx[bend] < sd[bend]) {
            cnt_ot += acc[bend];
            break;
          }
          cnt_ot -= acc[bend] * (sidx[bend] - 1);
          sidx[bend] = 0;
        }
      }
  }
  return 0;
}

void ora_gemm_rowmajor(const double* A, const double* B, double* C,
                       int64_t M, int64_t N, int64_t K) {
  for (int64_t i = 0; i < M; ++i) {
    double* c = C + i * N;
    for (int64_t j = 0; j < N; ++j) c[j] = 0.0;
    for (int64_t k = 0; k < K; ++k) {
      double a = A[i * K + k];
      const double* b = B + k * N;
      for (int64_t j = 0; j < N; ++j) c[j] += a * b[j];
    }
  }
}
