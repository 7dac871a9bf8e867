// Falm C++ host API: Region, IDX and Matrix<T> (src/region.h:8-31,
// src/util.h:29-31, src/matrix.h:12-327).
//
// Matrix<T> owns a host array and/or a device array of shape[0] rows (cells)
// by shape[1] columns, element (row, col) at ptr[row + col*shape[0]].  The
// device side is allocated through falm_malloc (zero-filled, as the
// reference's alloc does); the host side is zero-filled new[].
#ifndef FALM_AMD_MATRIX_H
#define FALM_AMD_MATRIX_H

#include <cstring>
#include <string>

#include "typedef.h"

namespace Falm {

static inline Int PRODUCT3(const Int3 &u) { return u[0] * u[1] * u[2]; }
static inline Int SUM3(const Int3 &u) { return u[0] + u[1] + u[2]; }
static inline Int IDX(Int i, Int j, Int k, const Int3 &shape) { return i + j * shape[0] + k * shape[0] * shape[1]; }

struct Region {
    Int3 shape;
    Int3 offset;
    Int size;
    Region() : shape(Int3{{0, 0, 0}}), offset(Int3{{0, 0, 0}}), size(0) {}
    Region(Int3 _shape, Int3 _offset) : shape(_shape), offset(_offset), size(PRODUCT3(_shape)) {}
    Region(Int3 _shape, Int gc)
        : Region(Int3{{_shape[0] - gc * 2, _shape[1] - gc * 2, _shape[2] - gc * 2}}, Int3{{gc, gc, gc}}) {}
    Region transform(Int3 st, Int3 ot) {
        return Region(Int3{{shape[0] + st[0], shape[1] + st[1], shape[2] + st[2]}},
                      Int3{{offset[0] + ot[0], offset[1] + ot[1], offset[2] + ot[2]}});
    }
};

enum class StencilMatrix { Empty, D3P7, D3P13, D2P5, D2P9, D1P3, D1P5 };

template <typename T>
struct MatrixFrame {
    T *ptr = nullptr;
    Int2 shape{{0, 0}};
    Int size = 0;
    Flag hdctype = HDC::Empty;
    StencilMatrix stencil = StencilMatrix::Empty;
    T &operator()(Int idx) const { return ptr[idx]; }
    T &operator()(Int row, Int col) const { return ptr[row + (size_t)col * shape[0]]; }

    void alloc(Int row, Int col, Flag hdc, StencilMatrix st = StencilMatrix::Empty) {
        shape = Int2{{row, col}};
        size = row * col;
        hdctype = hdc;
        stencil = st;
        if (hdc == HDC::Host) {
            ptr = new T[(size_t)size];
            std::memset(ptr, 0, sizeof(T) * (size_t)size);
        } else if (hdc == HDC::Device) {
            void *p = nullptr;
            falmErrCheckMacro(falm_malloc(&p, sizeof(T) * (size_t)size));
            ptr = (T *)p;
        }
    }
    void release() {
        if (hdctype == HDC::Host) delete[] ptr;
        else if (hdctype == HDC::Device) falmErrCheckMacro(falm_free(ptr));
        ptr = nullptr;
        hdctype = HDC::Empty;
    }
    void clear() {
        if (hdctype == HDC::Host) std::memset(ptr, 0, sizeof(T) * (size_t)size);
        else if (hdctype == HDC::Device) falmErrCheckMacro(falm_memset(ptr, 0, sizeof(T) * (size_t)size, nullptr));
    }
};

template <typename T>
struct Matrix {
    MatrixFrame<T> host, dev;
    Int2 shape{{0, 0}};
    Int size = 0;
    Flag hdctype = HDC::Empty;
    std::string name;
    StencilMatrix stencil = StencilMatrix::Empty;

    T &operator()(Int idx) const { return host(idx); }
    T &operator()(Int row, Int col) const { return host(row, col); }

    Matrix(const Matrix &) = delete;
    Matrix &operator=(const Matrix &) = delete;
    Matrix(std::string _name = "") : name(_name) {}
    Matrix(Int3 dom, Int dim, Flag hdc, const std::string &_name, StencilMatrix st = StencilMatrix::Empty) {
        alloc(dom, dim, hdc, _name, st);
    }
    Matrix(Int row, Int col, Flag hdc, const std::string &_name, StencilMatrix st = StencilMatrix::Empty) {
        alloc(row, col, hdc, _name, st);
    }
    ~Matrix() { release(); }

    void alloc(Int3 dom, Int dim, Flag hdc, const std::string &_name, StencilMatrix st = StencilMatrix::Empty) {
        alloc(PRODUCT3(dom), dim, hdc, _name, st);
    }
    void alloc(Int row, Int col, Flag hdc, const std::string &_name, StencilMatrix st = StencilMatrix::Empty) {
        if (hdctype != HDC::Empty) release();
        host.alloc(row, col, hdc & HDC::Host, st);
        dev.alloc(row, col, hdc & HDC::Device, st);
        shape = Int2{{row, col}};
        size = row * col;
        hdctype = hdc;
        name = _name;
        stencil = st;
    }
    void alloc(Int3 dom, Int dim, Flag hdc, StencilMatrix st = StencilMatrix::Empty) { alloc(dom, dim, hdc, name, st); }
    void alloc(Int row, Int col, Flag hdc, StencilMatrix st = StencilMatrix::Empty) { alloc(row, col, hdc, name, st); }

    void release(Flag hdc) {
        if (hdc & hdctype & HDC::Host) {
            host.release();
            hdctype &= ~HDC::Host;
        }
        if (hdc & hdctype & HDC::Device) {
            dev.release();
            hdctype &= ~HDC::Device;
        }
    }
    void release() {
        host.release();
        dev.release();
        hdctype = HDC::Empty;
    }
    // Matrix::sync (src/matrix.h:277-301): creates the missing side on demand
    void sync(Flag mcp) {
        const size_t bytes = sizeof(T) * (size_t)size;
        if (mcp == MCP::Hst2Dev) {
            if (!(hdctype & HDC::Device)) {
                dev.alloc(shape[0], shape[1], HDC::Device, stencil);
                hdctype |= HDC::Device;
            }
            falmErrCheckMacro(falm_memcpy(dev.ptr, host.ptr, bytes, MCP::Hst2Dev, nullptr));
        } else if (mcp == MCP::Dev2Hst) {
            if (!(hdctype & HDC::Host)) {
                host.alloc(shape[0], shape[1], HDC::Host, stencil);
                hdctype |= HDC::Host;
            }
            falmErrCheckMacro(falm_memcpy(host.ptr, dev.ptr, bytes, MCP::Dev2Hst, nullptr));
        }
    }
    void copy(Matrix<T> &src, Flag hdc) {
        const size_t bytes = sizeof(T) * (size_t)size;
        if (hdc & hdctype & HDC::Host) std::memcpy(host.ptr, src.host.ptr, bytes);
        if (hdc & hdctype & HDC::Device)
            falmErrCheckMacro(falm_memcpy(dev.ptr, src.dev.ptr, bytes, MCP::Dev2Dev, nullptr));
    }
    void clear(Flag hdc) {
        if (hdc & hdctype & HDC::Host) host.clear();
        if (hdc & hdctype & HDC::Device) dev.clear();
    }
    const char *cname() { return name.c_str(); }
};

template <typename T>
inline int falm_prec_of();
template <>
inline int falm_prec_of<double>() { return FALM_F64; }
template <>
inline int falm_prec_of<float>() { return FALM_F32; }

}  // namespace Falm

#endif
