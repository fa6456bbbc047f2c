// amrex_types.hpp -- the AMReX value types that TreESaP's plug-in signatures
// name, for programs that use this library without AMReX.
//
// TreESaP's concepts spell the right-hand side, boundary and initial-condition
// functors in AMReX types (pde-model/PDEModelConcepts.hpp:44-108):
// amrex::IntVect, amrex::Array4, amrex::GeometryData, amrex::BCRec, and the
// model holds an amrex::Geometry (pde-model/PDEModel.hpp:70-77). User RHS code
// written against TreESaP therefore names these types; this header supplies
// them with the same member functions and semantics for a single-level,
// single-box, cell-centred 3-D Cartesian domain, so such code compiles
// unchanged. Compiled by a HIP compiler, the function-space macros mark
// functions for host and device (functors the library does not carry run in a
// kernel compiled from them, device/generic.hpp); compiled by a host compiler
// they are empty.
//
// Semantics kept (AMReX's documented behaviour):
//   - Box is a cell-centred index box [smallEnd, bigEnd], inclusive;
//     grow(n) extends both ends by n (negative shrinks);
//   - Array4 is x-fastest, component-slowest, with its origin at `begin`;
//   - Geometry's cell size is (prob_hi - prob_lo) / n_cells per axis.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define AMREX_GPU_HOST_DEVICE __host__ __device__
#define AMREX_GPU_DEVICE __device__
#define AMREX_GPU_HOST __host__
#else
#define AMREX_GPU_HOST_DEVICE
#define AMREX_GPU_DEVICE
#define AMREX_GPU_HOST
#endif
#define AMREX_FORCE_INLINE inline
#define AMREX_SPACEDIM 3
#define AMREX_D_DECL(a, b, c) a, b, c

namespace amrex {

using Real = double;

template <class T> using Vector = std::vector<T>;
template <class T, std::size_t N> using Array = std::array<T, N>;

// Aggregate fixed-size array (brace-initialisable, trivially copyable).
template <class T, int N> struct GpuArray {
    T arr[N];
    constexpr T &operator[](int i) noexcept { return arr[i]; }
    constexpr const T &operator[](int i) const noexcept { return arr[i]; }
    constexpr T *data() noexcept { return arr; }
    constexpr const T *data() const noexcept { return arr; }
    static constexpr int size() noexcept { return N; }
};

template <class T> constexpr const T &Clamp(const T &v, const T &lo, const T &hi) {
    return (v < lo) ? lo : (hi < v) ? hi : v;
}

struct IntVect {
    int vect[3] = {0, 0, 0};

    constexpr IntVect() = default;
    constexpr IntVect(int i, int j, int k) : vect{i, j, k} {}
    constexpr explicit IntVect(int v) : vect{v, v, v} {}

    constexpr int &operator[](int d) noexcept { return vect[d]; }
    constexpr const int &operator[](int d) const noexcept { return vect[d]; }

    constexpr IntVect operator+(const IntVect &o) const { return {vect[0] + o[0], vect[1] + o[1], vect[2] + o[2]}; }
    constexpr IntVect operator-(const IntVect &o) const { return {vect[0] - o[0], vect[1] - o[1], vect[2] - o[2]}; }
    constexpr IntVect operator-() const { return {-vect[0], -vect[1], -vect[2]}; }
    constexpr IntVect operator*(int s) const { return {vect[0] * s, vect[1] * s, vect[2] * s}; }
    constexpr IntVect &operator+=(const IntVect &o) {
        for (int d = 0; d < 3; ++d) vect[d] += o[d];
        return *this;
    }
    constexpr IntVect &operator-=(const IntVect &o) {
        for (int d = 0; d < 3; ++d) vect[d] -= o[d];
        return *this;
    }
    constexpr bool operator==(const IntVect &o) const {
        return vect[0] == o[0] && vect[1] == o[1] && vect[2] == o[2];
    }
    constexpr bool operator!=(const IntVect &o) const { return !(*this == o); }
    constexpr bool allLE(const IntVect &o) const { return vect[0] <= o[0] && vect[1] <= o[1] && vect[2] <= o[2]; }
    constexpr bool allGE(const IntVect &o) const { return vect[0] >= o[0] && vect[1] >= o[1] && vect[2] >= o[2]; }

    static constexpr IntVect TheZeroVector() { return {0, 0, 0}; }
    static constexpr IntVect TheUnitVector() { return {1, 1, 1}; }
};

struct Dim3 {
    int x, y, z;
};

// Cell-centred index box, both ends inclusive.
class Box {
  public:
    constexpr Box() : lo_(0, 0, 0), hi_(-1, -1, -1) {}
    constexpr Box(const IntVect &small, const IntVect &big) : lo_(small), hi_(big) {}

    constexpr const IntVect &smallEnd() const { return lo_; }
    constexpr const IntVect &bigEnd() const { return hi_; }
    constexpr int smallEnd(int d) const { return lo_[d]; }
    constexpr int bigEnd(int d) const { return hi_[d]; }
    constexpr int length(int d) const { return hi_[d] - lo_[d] + 1; }
    constexpr IntVect length() const { return {length(0), length(1), length(2)}; }
    constexpr bool ok() const { return hi_.allGE(lo_); }
    constexpr std::int64_t numPts() const {
        return ok() ? (std::int64_t)length(0) * length(1) * length(2) : 0;
    }
    constexpr bool contains(const IntVect &p) const { return p.allGE(lo_) && p.allLE(hi_); }
    constexpr bool contains(int i, int j, int k) const { return contains(IntVect(i, j, k)); }
    constexpr Box &grow(int n) {
        for (int d = 0; d < 3; ++d) {
            lo_[d] -= n;
            hi_[d] += n;
        }
        return *this;
    }
    constexpr Box &grow(int d, int n) {
        lo_[d] -= n;
        hi_[d] += n;
        return *this;
    }
    constexpr bool operator==(const Box &o) const { return lo_ == o.lo_ && hi_ == o.hi_; }

  private:
    IntVect lo_, hi_;
};

constexpr Box grow(Box b, int n) { return b.grow(n); }

// Non-owning 4-D view: element (i, j, k, n) at p[(i-begin.x) + (j-begin.y)*jstride
// + (k-begin.z)*kstride + n*nstride].
template <class T> struct Array4 {
    T *p = nullptr;
    std::int64_t jstride = 0, kstride = 0, nstride = 0;
    Dim3 begin{0, 0, 0};
    Dim3 end{0, 0, 0}; // one past the last cell
    int ncomp = 0;

    constexpr Array4() = default;
    constexpr Array4(T *ptr, Dim3 b, Dim3 e, int nc)
        : p(ptr), jstride((std::int64_t)(e.x - b.x)), kstride((std::int64_t)(e.x - b.x) * (e.y - b.y)),
          nstride((std::int64_t)(e.x - b.x) * (e.y - b.y) * (e.z - b.z)), begin(b), end(e), ncomp(nc) {}
    // a view with explicit strides (padded rows), p pointing at element (b.x, b.y, b.z, 0)
    constexpr Array4(T *ptr, std::int64_t js, std::int64_t ks, std::int64_t ns, Dim3 b, Dim3 e, int nc)
        : p(ptr), jstride(js), kstride(ks), nstride(ns), begin(b), end(e), ncomp(nc) {}
    // a read-only view of a writable one
    template <class U>
        requires(std::is_same_v<const U, T> && !std::is_same_v<U, T>)
    constexpr Array4(const Array4<U> &o)
        : p(o.p), jstride(o.jstride), kstride(o.kstride), nstride(o.nstride), begin(o.begin), end(o.end),
          ncomp(o.ncomp) {}

    constexpr T &operator()(int i, int j, int k, int n = 0) const noexcept {
        return p[(i - begin.x) + (j - begin.y) * jstride + (k - begin.z) * kstride + n * nstride];
    }
    constexpr T &operator()(const IntVect &iv, int n = 0) const noexcept { return (*this)(iv[0], iv[1], iv[2], n); }
    constexpr bool contains(int i, int j, int k) const {
        return i >= begin.x && i < end.x && j >= begin.y && j < end.y && k >= begin.z && k < end.z;
    }
    constexpr int nComp() const { return ncomp; }
};

struct RealBox {
    Real xlo[3] = {0.0, 0.0, 0.0};
    Real xhi[3] = {1.0, 1.0, 1.0};

    RealBox() = default;
    RealBox(const std::array<Real, 3> &lo, const std::array<Real, 3> &hi) {
        for (int d = 0; d < 3; ++d) {
            xlo[d] = lo[d];
            xhi[d] = hi[d];
        }
    }
    RealBox(Real x0, Real y0, Real z0, Real x1, Real y1, Real z1) : xlo{x0, y0, z0}, xhi{x1, y1, z1} {}
    Real lo(int d) const { return xlo[d]; }
    Real hi(int d) const { return xhi[d]; }
    Real length(int d) const { return xhi[d] - xlo[d]; }
};

struct CoordSys {
    enum CoordType { undef = -1, cartesian = 0, RZ = 1, SPHERICAL = 2 };
};

// Plain-data geometry passed to functors by value (trivially copyable).
struct GeometryData {
    Real prob_domain_lo[3];
    Real prob_domain_hi[3];
    Real dx[3];
    Box domain;
    int is_periodic[3];
    int coord;

    AMREX_GPU_HOST_DEVICE Real CellSize(int d) const noexcept { return dx[d]; }
    AMREX_GPU_HOST_DEVICE const Real *CellSize() const noexcept { return dx; }
    AMREX_GPU_HOST_DEVICE Real ProbLo(int d) const noexcept { return prob_domain_lo[d]; }
    AMREX_GPU_HOST_DEVICE Real ProbHi(int d) const noexcept { return prob_domain_hi[d]; }
    AMREX_GPU_HOST_DEVICE const Box &Domain() const noexcept { return domain; }
    AMREX_GPU_HOST_DEVICE int isPeriodic(int d) const noexcept { return is_periodic[d]; }
    AMREX_GPU_HOST_DEVICE int Coord() const noexcept { return coord; }
};

class Geometry {
  public:
    Geometry() = default;
    Geometry(const Box &dom, const RealBox *rb, int coord, int const *is_per) { define(dom, rb, coord, is_per); }
    Geometry(const Box &dom, const RealBox &rb, int coord, const std::array<int, 3> &is_per) {
        define(dom, &rb, coord, is_per.data());
    }

    void define(const Box &dom, const RealBox *rb, int coord, int const *is_per) {
        d_.domain = dom;
        d_.coord = coord;
        const RealBox r = rb ? *rb : RealBox{};
        for (int a = 0; a < 3; ++a) {
            d_.prob_domain_lo[a] = r.lo(a);
            d_.prob_domain_hi[a] = r.hi(a);
            d_.dx[a] = (r.hi(a) - r.lo(a)) / static_cast<Real>(dom.length(a));
            d_.is_periodic[a] = is_per ? (is_per[a] != 0) : 0;
        }
    }

    const Box &Domain() const { return d_.domain; }
    Real CellSize(int d) const { return d_.dx[d]; }
    GpuArray<Real, 3> CellSizeArray() const { return {{d_.dx[0], d_.dx[1], d_.dx[2]}}; }
    GpuArray<Real, 3> InvCellSizeArray() const { return {{1.0 / d_.dx[0], 1.0 / d_.dx[1], 1.0 / d_.dx[2]}}; }
    Real ProbLo(int d) const { return d_.prob_domain_lo[d]; }
    Real ProbHi(int d) const { return d_.prob_domain_hi[d]; }
    GpuArray<Real, 3> ProbLoArray() const { return {{d_.prob_domain_lo[0], d_.prob_domain_lo[1], d_.prob_domain_lo[2]}}; }
    GpuArray<Real, 3> ProbHiArray() const { return {{d_.prob_domain_hi[0], d_.prob_domain_hi[1], d_.prob_domain_hi[2]}}; }
    bool isPeriodic(int d) const { return d_.is_periodic[d] != 0; }
    bool isAllPeriodic() const { return isPeriodic(0) && isPeriodic(1) && isPeriodic(2); }
    int Coord() const { return d_.coord; }
    const GeometryData &data() const { return d_; }

  private:
    GeometryData d_{{0, 0, 0}, {1, 1, 1}, {1, 1, 1}, Box{}, {0, 0, 0}, 0};
};

// Boundary-condition record (per component, lo and hi side per axis). The
// physical-boundary functors of this path ignore it, as TreESaP's do.
struct BCType {
    enum mathematicalBndryTypes : int {
        bogus = -666,
        reflect_odd = -1,
        int_dir = 0,
        reflect_even = 1,
        foextrap = 2,
        ext_dir = 3,
        hoextrap = 4
    };
};

class BCRec {
  public:
    BCRec() = default;
    AMREX_GPU_HOST_DEVICE BCRec(int loX, int loY, int loZ, int hiX, int hiY, int hiZ)
        : bc{loX, loY, loZ, hiX, hiY, hiZ} {}
    AMREX_GPU_HOST_DEVICE int lo(int d) const { return bc[d]; }
    AMREX_GPU_HOST_DEVICE int hi(int d) const { return bc[3 + d]; }
    AMREX_GPU_HOST_DEVICE void setLo(int d, int v) { bc[d] = v; }
    AMREX_GPU_HOST_DEVICE void setHi(int d, int v) { bc[3 + d] = v; }

  private:
    int bc[6] = {BCType::bogus, BCType::bogus, BCType::bogus, BCType::bogus, BCType::bogus, BCType::bogus};
};

// Process lifetime hooks. The library initialises the device lazily on the
// first solver, so these only exist for source compatibility.
inline void Initialize(int &, char **&) {}
inline void Finalize() {}

} // namespace amrex
