// Falm C++ host API over the falm_amd C ABI: basic types and flags.
//
// Mirrors src/typedef.h:14-41, src/vectypes.h, src/flag.h:8-29 and
// src/error.h:9-30 of the reference so that code written against the
// reference's class API compiles against this one.  CUDA types are replaced
// by plain ones: dim3 is a launch-shape hint (ignored: the library chooses
// its own tiling) and Stream is an opaque hipStream_t.
#ifndef FALM_AMD_TYPEDEF_H
#define FALM_AMD_TYPEDEF_H

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../falm_amd.h"

namespace Falm {

typedef double Real;
typedef unsigned Flag;
typedef int Int;
typedef void *Stream;

template <typename T, int N>
struct Vector {
    T _v[N];
    T &operator[](int i) { return _v[i]; }
    const T &operator[](int i) const { return _v[i]; }
    Vector operator+(const Vector &o) const {
        Vector r;
        for (int i = 0; i < N; i++) r._v[i] = _v[i] + o._v[i];
        return r;
    }
    Vector operator-(const Vector &o) const {
        Vector r;
        for (int i = 0; i < N; i++) r._v[i] = _v[i] - o._v[i];
        return r;
    }
    Vector operator-(T s) const {
        Vector r;
        for (int i = 0; i < N; i++) r._v[i] = _v[i] - s;
        return r;
    }
    Vector operator+(T s) const {
        Vector r;
        for (int i = 0; i < N; i++) r._v[i] = _v[i] + s;
        return r;
    }
    Vector &operator+=(T s) {
        for (int i = 0; i < N; i++) _v[i] += s;
        return *this;
    }
    Vector &operator-=(T s) {
        for (int i = 0; i < N; i++) _v[i] -= s;
        return *this;
    }
};

typedef Vector<Int, 3> Int3;
typedef Vector<Real, 3> Real3;
typedef Vector<Int, 2> Int2;

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

const Int GuideCell = 2;

// src/flag.h:8-29
class MCP {
public:
    static const Flag Empty = 0;
    static const Flag Hst2Hst = 1;
    static const Flag Hst2Dev = 2;
    static const Flag Dev2Hst = 4;
    static const Flag Dev2Dev = 8;
};
class HDC {
public:
    static const Flag Empty = 0;
    static const Flag Host = 1;
    static const Flag Device = 2;
    static const Flag HstDev = Host | Device;
};
class Color {
public:
    static const Int Black = 0;
    static const Int Red = 1;
};

// src/error.h:9-18 (extended with the C ABI's codes)
class FalmErr {
public:
    static const int success = FALM_OK;
    static const int mallocErr = FALM_ERR_MALLOC;
    static const int setUpFileErr = FALM_ERR_SETUP;
    static const int cuErrMask = 100;
};

// The reference prints failing calls and continues (falmErrCheckMacro,
// src/error.h:20-26).  The same happens here; defining
// FALM_THROW_ON_ERROR turns a failure into a std::runtime_error.
inline int &lastStatus() {
    static thread_local int s = FALM_OK;
    return s;
}

inline int FalmErrCheck(int errCode, const char *name, const char *file, int line) {
    if (errCode != FALM_OK) {
        lastStatus() = errCode;
        fprintf(stderr, "%s at %s %d failed with %d (%s: %s)\n", name, file, line, errCode,
                falm_status_string(errCode), falm_last_error());
#ifdef FALM_THROW_ON_ERROR
        throw std::runtime_error(std::string(name) + ": " + falm_last_error());
#endif
    }
    return errCode;
}

}  // namespace Falm

#define falmErrCheckMacro(expr) Falm::FalmErrCheck((expr), #expr, __FILE__, __LINE__)

#endif
