// Tensor<Ty>: the reference's dense container (include/tensor.hpp:44-182,
// template/tensor_mpi.tpp).  Host-addressable column-major storage shared on
// copy/assignment (the reference's static ref-count map, tensor_mpi.tpp:17-18,
// is a shared_ptr here), local == global shape at one rank.  The GPU holds
// mirrors only for the duration of a Function/Algorithm call.
#ifndef DIANA_AMD_TENSOR_HPP
#define DIANA_AMD_TENSOR_HPP
#include <cassert>
#include <cstdarg>
#include <iostream>
#include <memory>
#include <string>
#include "def.hpp"
#include "communicator.hpp"
#include "distribution.hpp"
#include "operator.hpp"

template <typename Ty>
class Tensor {
    std::shared_ptr<Ty> buf_;
    Ty *data_ = nullptr;
    shape_t shape_, stride_;
    size_t ndim_ = 0, size_ = 0;
    Distribution *distribution_ = nullptr;
    Operator<Ty> op_;

    void init_by_shape(const shape_t &shape) {
        shape_ = shape;
        ndim_ = shape.size();
        size_ = 1;
        stride_.assign(ndim_, 1);
        for (size_t i = 0; i < ndim_; i++) {
            size_ *= shape[i];
            if (i > 0) stride_[i] = stride_[i - 1] * shape[i - 1];
        }
    }
    void allocate(bool zero) {
        buf_ = std::shared_ptr<Ty>((Ty *)std::calloc(size_ ? size_ : 1, sizeof(Ty)), [](Ty *p) { std::free(p); });
        data_ = buf_.get();
        (void)zero;  // calloc already zeroes
    }

public:
    Tensor() = default;
    Tensor(Ty *A, const shape_t &shape) { init_by_shape(shape); data_ = A; }  // external, not owned
    explicit Tensor(const shape_t &shape, bool zero = true) { init_by_shape(shape); allocate(zero); }
    Tensor(Distribution *d, Ty *A, const shape_t &shape) : distribution_(d) { init_by_shape(shape); data_ = A; }
    Tensor(Distribution *d, const shape_t &shape, bool zero = true) : distribution_(d) {
        init_by_shape(shape);
        allocate(zero);
    }

    Ty *data() const { return data_; }
    Operator<Ty> *op() const { return const_cast<Operator<Ty> *>(&op_); }
    size_t ndim() const { return ndim_; }
    size_t size() const { return size_; }
    const shape_t &shape() const { return shape_; }
    const shape_t &stride() const { return stride_; }
    const shape_t &shape_global() const { return shape_; }
    size_t size_global() const { return size_; }
    bool is_matrix() const { return ndim_ == 2; }
    Distribution *distribution() const { return distribution_; }
    int comm_rank() const { return 0; }
    int comm_size() const { return 1; }
    void sync(int) {}  // MPI_Bcast from one rank to itself (tensor_mpi.tpp:403-406)

    Ty &operator[](size_t i) { return data_[i]; }
    const Ty &operator[](size_t i) const { return data_[i]; }
    Ty &operator()(size_t x, ...) {
        size_t index = x * stride_[0];
        va_list args;
        va_start(args, x);
        for (size_t i = 1; i < ndim_; i++) index += va_arg(args, size_t) * stride_[i];
        va_end(args);
        return data_[index];
    }
    void constant(Ty c) { op_.constant(data_, c, size_); }
    void zeros() { constant(0); }
    void ones() { constant(1); }
    void rand() { op_.rand(data_, size_); }
    void randn() { op_.randn(data_, size_); }
    double fnorm() const { return Operator<Ty>::fnorm(data_, size_); }
    Ty sum() const { return Operator<Ty>::sum(data_, size_); }
    Tensor<Ty> copy() const {
        Tensor<Ty> r = distribution_ ? Tensor<Ty>(distribution_, shape_, false) : Tensor<Ty>(shape_, false);
        std::memcpy(r.data_, data_, size_ * sizeof(Ty));
        return r;
    }
    void print() const {
        std::string s = "[";
        for (size_t i = 0; i < size_; i++) s += std::to_string(data_[i]) + (i + 1 < size_ ? " " : "");
        std::cerr << s << "]" << std::endl;
    }
};
#endif
