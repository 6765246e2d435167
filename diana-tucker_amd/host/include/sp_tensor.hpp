// SpTensor (include/sp_tensor.hpp:26-240) over libdiana_hip.so.  The
// constructor keeps the reference's host-visible result -- index lists and
// values reordered by a stable sort on the slice mode (the largest mode,
// first on ties; sp_tensor.hpp:107-147) -- and builds the device copy
// (dt_sp_create: int32 indices, FP32 values, one stably sorted copy per
// mode).  The permutation comes from the device sort, which is the
// counting sort's.  One rank: no slice distribution.
#ifndef DIANA_AMD_SP_TENSOR_HPP
#define DIANA_AMD_SP_TENSOR_HPP
#include <cstdlib>
#include <vector>
#include "device.hpp"
#include "logger.hpp"
#include "read_data.hpp"

template <typename Ty>
class SpTensor {
    size_t **index_lists_ = nullptr;
    Ty *vals_ = nullptr;
    shape_t dims_;
    size_t ndim_ = 0, nnz_ = 0;
    dt_sp *sp_ = nullptr;

public:
    size_t slice_mode = 0;
    size_t slice_start = 0, slice_end = 0;
    shape_t send_to_list, recv_from_list;

    SpTensor() = default;
    explicit SpTensor(data_buffer<Ty> *dbf) {
        ndim_ = dbf->ndim;
        nnz_ = dbf->nnz;
        for (size_t n = 0; n < ndim_; n++) dims_.push_back(dbf->dims[n]);
        size_t best = 0;
        for (size_t n = 0; n < ndim_; n++)
            if (best < dims_[n]) {
                best = dims_[n];
                slice_mode = n;
            }
        slice_start = 0;
        slice_end = ndim_ ? dims_[slice_mode] - 1 : 0;
        if (nnz_ >= ((size_t)1 << 31)) error("SpTensor: nnz must stay below 2^31");
        dt_ctx *ctx = diana::Device::ctx();
        std::vector<int32_t> i32(nnz_);
        std::vector<float> v32(nnz_);
        std::vector<diana::DevBuf *> di;
        std::vector<const int32_t *> dip;
        for (size_t n = 0; n < ndim_; n++) {
            for (size_t k = 0; k < nnz_; k++) i32[k] = (int32_t)dbf->index_lists[n][k];
            di.push_back(new diana::DevBuf(i32.data(), nnz_ * 4));
            dip.push_back((const int32_t *)di.back()->p);
        }
        for (size_t k = 0; k < nnz_; k++) v32[k] = (float)dbf->vals[k];
        diana::DevBuf dv(v32.data(), nnz_ * 4);
        size_t dd[DT_MAX_NDIM];
        for (size_t n = 0; n < ndim_; n++) dd[n] = dims_[n];
        diana::check(dt_sp_create(ctx, (int)ndim_, dd, (int64_t)nnz_, dip.data(), (const float *)dv.p, &sp_),
                     "SpTensor");
        for (auto *b : di) delete b;
        // host arrays in slice-mode order
        const int32_t *dperm = nullptr;
        diana::check(dt_sp_mode_arrays(sp_, (int)slice_mode, &dperm, nullptr, nullptr), "SpTensor");
        std::vector<int32_t> perm(nnz_);
        diana::check(dt_d2h(ctx, perm.data(), dperm, nnz_ * 4), "SpTensor");
        index_lists_ = (size_t **)std::malloc(sizeof(size_t *) * ndim_);
        for (size_t n = 0; n < ndim_; n++) {
            index_lists_[n] = (size_t *)std::malloc(sizeof(size_t) * (nnz_ ? nnz_ : 1));
            for (size_t k = 0; k < nnz_; k++) index_lists_[n][k] = dbf->index_lists[n][perm[k]];
        }
        vals_ = (Ty *)std::malloc(sizeof(Ty) * (nnz_ ? nnz_ : 1));
        for (size_t k = 0; k < nnz_; k++) vals_[k] = dbf->vals[perm[k]];
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
};
#endif
