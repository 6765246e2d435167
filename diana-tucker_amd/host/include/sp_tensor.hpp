// SpTensor (include/sp_tensor.hpp:26-240) over libdiana_hip.so.  The
// constructor keeps the reference's host-visible result -- index lists and
// values reordered by a stable sort on the slice mode (the largest mode,
// first on ties; sp_tensor.hpp:107-147) -- and builds the device copy
// (dt_sp_create: int32 indices, one stably sorted copy per mode).  The
// permutation comes from the device sort, which is the counting sort's.  One
// rank: no slice distribution.
//
// Two ingest choices the reference does not have (diana::sp_options()):
//   merge_duplicates (default on): nonzeros with equal coordinates are merged
//     first (dt_sp_merge_duplicates: values summed in FP64 at the first
//     occurrence, input order kept), so a file without duplicates is taken
//     as it is;
//   fp32_values (default off): SpTensor<double> keeps FP64 values on the
//     device (dt_sp_create_f64: TTMc and factor updates in FP64, as the
//     reference computes); set it to narrow them to FP32 and take the FP32
//     kernels (3-way rank <= 16: the MFMA TTMc).  SpTensor<float> is FP32.
#ifndef DIANA_AMD_SP_TENSOR_HPP
#define DIANA_AMD_SP_TENSOR_HPP
#include <cstdlib>
#include <type_traits>
#include <vector>
#include "device.hpp"
#include "logger.hpp"
#include "read_data.hpp"

namespace diana {
struct SpOptions {
    bool merge_duplicates = true;
    bool fp32_values = false;
};
inline SpOptions &sp_options() {
    static SpOptions o;
    return o;
}
}  // namespace diana

template <typename Ty>
class SpTensor {
    size_t **index_lists_ = nullptr;
    Ty *vals_ = nullptr;
    shape_t dims_;
    size_t ndim_ = 0, nnz_ = 0;
    dt_sp *sp_ = nullptr;
    bool fp64_ = false;

public:
    size_t slice_mode = 0;
    size_t slice_start = 0, slice_end = 0;
    shape_t send_to_list, recv_from_list;

    SpTensor() = default;
    explicit SpTensor(data_buffer<Ty> *dbf) {
        ndim_ = dbf->ndim;
        nnz_ = dbf->nnz;
        if (ndim_ < 2 || ndim_ > DT_MAX_NDIM) error("SpTensor: " + std::to_string(ndim_) + " modes (2 to 8 supported)");
        if (nnz_ == 0) error("SpTensor: no nonzeros");
        for (size_t n = 0; n < ndim_; n++) dims_.push_back(dbf->dims[n]);
        size_t best = 0;
        for (size_t n = 0; n < ndim_; n++)
            if (best < dims_[n]) {
                best = dims_[n];
                slice_mode = n;
            }
        slice_start = 0;
        slice_end = dims_[slice_mode] - 1;
        if (nnz_ >= ((size_t)1 << 31)) error("SpTensor: nnz must stay below 2^31");
        for (size_t n = 0; n < ndim_; n++)
            if (dims_[n] >= ((size_t)1 << 31)) error("SpTensor: mode sizes must stay below 2^31");
        fp64_ = std::is_same<Ty, double>::value && !diana::sp_options().fp32_values;
        dt_ctx *ctx = diana::Device::ctx();
        // device copies: int32 indices, FP64 values (merged in FP64)
        std::vector<int32_t> i32(nnz_);
        std::vector<diana::DevBuf *> di;
        std::vector<int32_t *> dip;
        for (size_t n = 0; n < ndim_; n++) {
            for (size_t k = 0; k < nnz_; k++) i32[k] = (int32_t)dbf->index_lists[n][k];
            di.push_back(new diana::DevBuf(i32.data(), nnz_ * 4));
            dip.push_back((int32_t *)di.back()->p);
        }
        std::vector<double> v64(nnz_);
        for (size_t k = 0; k < nnz_; k++) v64[k] = (double)dbf->vals[k];
        diana::DevBuf dv(v64.data(), nnz_ * 8);
        size_t dd[DT_MAX_NDIM];
        for (size_t n = 0; n < ndim_; n++) dd[n] = dims_[n];
        // host lists of the (merged) tensor in input order
        std::vector<std::vector<size_t>> lists(ndim_);
        std::vector<double> vals_in;
        if (diana::sp_options().merge_duplicates) {
            int64_t nu = 0;
            diana::check(dt_sp_merge_duplicates(ctx, (int)ndim_, dd, (int64_t)nnz_, dip.data(), (double *)dv.p, &nu),
                         "SpTensor");
            nnz_ = (size_t)nu;
            for (size_t n = 0; n < ndim_; n++) {
                diana::check(dt_d2h(ctx, i32.data(), dip[n], nnz_ * 4), "SpTensor");
                lists[n].assign(i32.begin(), i32.begin() + nnz_);
            }
            vals_in.resize(nnz_);
            diana::check(dt_d2h(ctx, vals_in.data(), dv.p, nnz_ * 8), "SpTensor");
        } else {
            for (size_t n = 0; n < ndim_; n++) lists[n].assign(dbf->index_lists[n], dbf->index_lists[n] + nnz_);
            vals_in = v64;
        }
        std::vector<const int32_t *> cip(dip.begin(), dip.end());
        if (fp64_) {
            diana::check(dt_sp_create_f64(ctx, (int)ndim_, dd, (int64_t)nnz_, cip.data(), (const double *)dv.p, &sp_),
                         "SpTensor");
        } else {
            std::vector<float> v32(nnz_);
            for (size_t k = 0; k < nnz_; k++) v32[k] = (float)vals_in[k];
            diana::DevBuf dv32(v32.data(), nnz_ * 4);
            diana::check(dt_sp_create(ctx, (int)ndim_, dd, (int64_t)nnz_, cip.data(), (const float *)dv32.p, &sp_),
                         "SpTensor");
        }
        for (auto *b : di) delete b;
        // host arrays in slice-mode order
        const int32_t *dperm = nullptr;
        diana::check(dt_sp_mode_arrays(sp_, (int)slice_mode, &dperm, nullptr, nullptr), "SpTensor");
        std::vector<int32_t> perm(nnz_);
        diana::check(dt_d2h(ctx, perm.data(), dperm, nnz_ * 4), "SpTensor");
        index_lists_ = (size_t **)std::malloc(sizeof(size_t *) * ndim_);
        for (size_t n = 0; n < ndim_; n++) {
            index_lists_[n] = (size_t *)std::malloc(sizeof(size_t) * nnz_);
            for (size_t k = 0; k < nnz_; k++) index_lists_[n][k] = lists[n][perm[k]];
        }
        vals_ = (Ty *)std::malloc(sizeof(Ty) * nnz_);
        for (size_t k = 0; k < nnz_; k++) vals_[k] = (Ty)vals_in[perm[k]];
    }
    SpTensor(const SpTensor &) = delete;
    SpTensor &operator=(const SpTensor &) = delete;
    ~SpTensor() {
        if (sp_) dt_sp_destroy(sp_);
        if (index_lists_) {
            for (size_t n = 0; n < ndim_; n++) std::free(index_lists_[n]);
            std::free(index_lists_);
        }
        std::free(vals_);
    }
    size_t ndim() const { return ndim_; }
    size_t nnz() const { return nnz_; }
    const shape_t &shape() const { return dims_; }
    size_t **index_lists() const { return index_lists_; }
    const Ty *vals() const { return vals_; }
    dt_sp *device() const { return sp_; }
    bool fp64() const { return fp64_; }
};
#endif
