// Kept host API: device and pinned-host arrays with real and complex views of
// one buffer (cuda/array_abstraction.hpp:15-212), on libmcstar_hip.so.
//
//   real view    (i2*ny + i1)*nx + i0, unpadded
//   complex view (i2*ny + i1)*(nx/2+1) + i0
//   bytes        (nx/2+1)*ny*nz*2*sizeof(F)
// Copies (>>) are synchronous; % swaps buffers; arrays own their memory and
// are not copyable.  Errors fail fast (print, exit) as CUERR did.
#pragma once

#include <hip/hip_runtime.h>

#include <cassert>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../common/cu_error.hpp"
#include "../common/npyutil.hpp"
#include "../common/technicalities.hpp"

namespace Array {

struct AsReal {};
struct AsCmpl {};

// type provider (array_abstraction.hpp:21-25)
template <typename> struct TP {};
template <> struct TP<float> { typedef float RT; typedef float2 CT; };
template <> struct TP<double> { typedef double RT; typedef double2 CT; };

template <typename F> constexpr int prec_of() { return sizeof(F) == 8 ? MCS_F64 : MCS_F32; }

template <typename F, typename Kind> struct NpyWriter;
template <typename F> struct NpyWriter<F, AsReal> {
    static void save(const std::string& fn, int dims, const int* rshape, const int*, const void* p) {
        aoba::SaveArrayAsNumpy(fn, dims, rshape, reinterpret_cast<const F*>(p));
    }
};
template <typename F> struct NpyWriter<F, AsCmpl> {
    static void save(const std::string& fn, int dims, const int*, const int* cshape, const void* p) {
        aoba::SaveArrayAsNumpy(fn, dims, cshape, reinterpret_cast<const std::complex<F>*>(p));
    }
};

template <typename F> class Array {
  protected:
    F* m_ptr;
    const int m_dims;
    int m_real_elems;
    int3 m_ext_real;
    int3 m_ext_cmpl;
    int m_cmpl_elems;
    size_t m_bytes;
    explicit Array(int3 ext)
        : m_ptr(nullptr),
          m_dims(ext.z == 1 ? (ext.y == 1 ? 1 : 2) : 3),
          m_real_elems(ext.x * ext.y * ext.z),
          m_ext_real(ext),
          m_ext_cmpl({ext.x / 2 + 1, ext.y, ext.z}),
          m_cmpl_elems((ext.x / 2 + 1) * ext.y * ext.z),
          m_bytes((size_t)(ext.x / 2 + 1) * ext.y * ext.z * 2 * sizeof(F)) {
        if (ext.x < 1 || ext.y < 1 || ext.z < 1) throw std::runtime_error("extent must be >=1 in all dimensions");
    }
    Array(const Array&) = delete;
    Array& operator=(const Array&) = delete;

  public:
    int real_elems() { return m_real_elems; }
    int cmpl_elems() { return m_cmpl_elems; }
    int3 real_vext() { return m_ext_real; }
    int3 cmpl_vext() { return m_ext_cmpl; }
    int real_ext(int i) { return i == 0 ? m_ext_real.x : (i == 1 ? m_ext_real.y : m_ext_real.z); }
    int cmpl_ext(int i) { return i == 0 ? m_ext_cmpl.x : (i == 1 ? m_ext_cmpl.y : m_ext_cmpl.z); }
    size_t bytes() const { return m_bytes; }
    void* ptr_void() { return m_ptr; }
    F* ptr_real() { return m_ptr; }
    typename TP<F>::CT* ptr_cmpl() { return reinterpret_cast<typename TP<F>::CT*>(m_ptr); }
    const int* ext_ptr() const { return reinterpret_cast<const int*>(&m_ext_real); }

    typedef F RT;
    typedef typename TP<F>::CT CT;
};

template <typename> struct CPUArray;

template <typename F> struct GPUArray : public Array<F> {
    explicit GPUArray(CPUArray<F>& from) : Array<F>(from.real_vext()) {
        CUERR(mcs_malloc((void**)&this->m_ptr, this->m_bytes));
        from >> (*this);
    }
    explicit GPUArray(int3 ext) : Array<F>(ext) { CUERR(mcs_malloc((void**)&this->m_ptr, this->m_bytes)); }
    ~GPUArray() { CUERR(mcs_free(this->m_ptr)); }
    void operator>>(GPUArray<F>& o) {
        assert(o.real_vext() == this->m_ext_real && "array dimensions differ");
        CUERR(mcs_memcpy(o.ptr_void(), this->m_ptr, this->m_bytes, MCS_D2D));
    }
    void operator>>(CPUArray<F>& o) {
        assert(o.real_vext() == this->m_ext_real && "array dimensions differ");
        CUERR(mcs_memcpy(o.ptr_void(), this->m_ptr, this->m_bytes, MCS_D2H));
    }
    void operator%(GPUArray<F>& o) {
        assert(o.real_vext() == this->m_ext_real && "array dimensions differ");
        F* p = o.m_ptr;
        o.m_ptr = this->m_ptr;
        this->m_ptr = p;
    }
};

template <typename F> struct CPUArray : public Array<F> {
    explicit CPUArray(CPUArray<F>& from) : Array<F>(from.real_vext()) {
        CUERR(mcs_host_alloc((void**)&this->m_ptr, this->m_bytes));
        from >> (*this);
    }
    explicit CPUArray(int3 ext) : Array<F>(ext) {
        CUERR(mcs_host_alloc((void**)&this->m_ptr, this->m_bytes));  // zero-filled
    }
    ~CPUArray() { CUERR(mcs_host_free(this->m_ptr)); }
    void operator>>(CPUArray<F>& o) {
        assert(o.real_vext() == this->m_ext_real && "array dimensions differ");
        std::memcpy(o.ptr_void(), this->m_ptr, this->m_bytes);
    }
    void operator>>(GPUArray<F>& o) {
        assert(o.real_vext() == this->m_ext_real && "array dimensions differ");
        CUERR(mcs_memcpy(o.ptr_void(), this->m_ptr, this->m_bytes, MCS_H2D));
    }
    void operator%(CPUArray<F>& o) {
        assert(o.real_vext() == this->m_ext_real && "array dimensions differ");
        F* p = o.m_ptr;
        o.m_ptr = this->m_ptr;
        this->m_ptr = p;
    }
    // idx.x fastest (array_abstraction.hpp:133-136)
    inline F& operator[](int idx) { return this->m_ptr[idx]; }
    inline F& operator[](int3 idx) {
        return this->m_ptr[(idx.z * this->m_ext_real.y + idx.y) * this->m_ext_real.x + idx.x];
    }
    // npy load: numpy shape (z, y, x); a mismatch aborts (array_abstraction.hpp:138-153)
    void load(const std::string fn) {
        std::vector<F> data;
        std::vector<int> shape;
        aoba::LoadArrayFromNumpy(fn, shape, data);
        int3 s{1, 1, 1};
        switch (shape.size()) {
            case 1: s.x = shape[0]; break;
            case 2: s.x = shape[1]; s.y = shape[0]; break;
            case 3: s.x = shape[2]; s.y = shape[1]; s.z = shape[0]; break;
            default: throw std::runtime_error("npy: unsupported rank");
        }
        if (!(s == this->m_ext_real)) {
            fprintf(stderr, "*** array dimension mismatch");
            fflush(stderr);
            abort();
        }
        std::memcpy(this->m_ptr, data.data(), (size_t)this->m_real_elems * sizeof(F));
        fprintf(stderr, "*** array initialized from file <%s>\n", fn.c_str());
        fflush(stderr);
    }
    // npy save of the real view (numpy shape (z, y, x)) or the complex half spectrum
    template <typename Kind = AsReal> void save(const std::string fn) {
        int3 rshape = shape_tr(this->real_vext());
        int3 cshape = shape_tr(this->cmpl_vext());
        NpyWriter<F, Kind>::save(fn, this->m_dims, (int*)&rshape, (int*)&cshape, this->m_ptr);
    }
    void over(std::function<void(F&)> f) {
        for (int i = 0; i < this->m_real_elems; ++i) f(this->m_ptr[i]);
    }
    void over(std::function<void(typename TP<F>::CT&)> f) {
        for (int i = 0; i < this->m_cmpl_elems; ++i) f(this->ptr_cmpl()[i]);
    }
    void over(std::function<void(F&, int)> f) {
        for (int i = 0; i < this->m_real_elems; ++i) f(this->m_ptr[i], i);
    }
    void over(std::function<void(typename TP<F>::CT&, int)> f) {
        for (int i = 0; i < this->m_cmpl_elems; ++i) f(this->ptr_cmpl()[i], i);
    }
    void square() { over([](F& v) { v *= v; }); }
    void set_to(F val) { over([val](F& v) { v = val; }); }
    void mult(F val) { over([val](F& v) { v *= val; }); }
    void add(F val) { over([val](F& v) { v += val; }); }
    F max() {
        F m = this->m_ptr[0];
        over([&m](F& v) { m = m > v ? m : v; });
        return m;
    }
    F min() {
        F m = this->m_ptr[0];
        over([&m](F& v) { m = m < v ? m : v; });
        return m;
    }
};

}  // namespace Array
