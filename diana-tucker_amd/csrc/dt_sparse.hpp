// Internal layout of the sparse tensor (dt_sp) shared by the sparse sources.
#pragma once
#include "dt_internal.hpp"

typedef int dt_intx4 __attribute__((ext_vector_type(4)));

// One mode's stably sorted copy (the SpTensor constructor's counting sort,
// include/sp_tensor.hpp:117-147, applied to mode n).
struct dt_sp_mode {
    int32_t *perm = nullptr;    // sorted position -> input position
    int32_t *row = nullptr;     // sorted mode-n index of each nonzero
    int32_t *rowptr = nullptr;  // I_n + 1 row pointers
    int32_t *slot = nullptr;    // row -> compact number of the non-empty rows
    int32_t *rows = nullptr;    // compact number -> row
    dt_intx4 *rec = nullptr;    // 3-way tensors: TTMc records in this mode's order
    int64_t nrows = 0;          // non-empty rows
};

struct dt_sp {
    int ndim = 0;
    int64_t dims[DT_MAX_NDIM] = {};
    int64_t nnz = 0;
    int slice_mode = 0;
    int32_t *idx[DT_MAX_NDIM] = {};  // input order
    float *val = nullptr;            // input order, FP32
    double *val64 = nullptr;         // input order, FP64 (tensors created with FP64 values only)
    dt_sp_mode mode[DT_MAX_NDIM];
    double normA2 = 0.0;
    std::vector<void *> owned;
};

// ---- any number of modes and ranks (dt_sparse_nway.hip) ---------------------
// Element type T of the factor rows, Y_(n) and accumulation: DT_F32 (FP32
// values) or DT_F64 (needs val64).  Urow[m]: I_m x R_m row-major (a factor row
// is contiguous, the reference's Ut layout).  K = prod_{m != n} R_m <= DT_SPG_KMAX.
constexpr int DT_SPG_KMAX = 4096, DT_SPG_RMAX = 64;
// Y_c (nrows x K row-major, compact rows of mode n): sum over each row's
// nonzeros of val * (x)_{m != n} Urow[m][i_m, :], the other modes ascending
// with the first fastest (add_outer_product_all's order)
int dt_spg_ttmc(dt_ctx *ctx, const dt_sp *sp, int n, const void *const *Urow, int T, const int *R, void *Yc);
// L (K x R_n column-major FP64) = Y_cᵀ Urow[n][rows, :]
int dt_spg_ytu(dt_ctx *ctx, const dt_sp *sp, int n, const void *Un_row, int T, const int *R, const void *Yc,
               double *L);
// W (I_n x R_n column-major FP64, zero rows for empty slices) = Y_(n) L
int dt_spg_yl(dt_ctx *ctx, const dt_sp *sp, int n, int T, const int *R, const void *Yc, const double *L, double *W);
// Urow (I x R row-major, T) from U64 (I x R column-major FP64)
int dt_spg_rows_view(dt_ctx *ctx, const double *U64, int64_t I, int R, void *Urow, int T);
// ||v||_2 of n FP64 values into out (device)
int dt_spg_norm(dt_ctx *ctx, const double *v, int64_t n, double *out);
