// fields.hip -- Minkowski Yee stencils, current source term, digital filter,
// periodic ghost fill and periodic current fold.
//
// One thread per active cell, x fastest (64-wide rows), so that every
// wavefront reads and writes contiguous 256-byte segments of each component
// plane.  These kernels are HBM-bound.
#include "ntt_common.hpp"

namespace ntt {
namespace {

constexpr int BX = 64, BY = 4;

struct CellIdx {
  int i, j, k;
  bool ok;
};

__device__ __forceinline__ CellIdx active_cell(const FieldView& F) {
  CellIdx c;
  const int a = blockIdx.x * BX + threadIdx.x;
  const int b = blockIdx.y * BY + threadIdx.y;
  const int g = blockIdx.z;
  c.ok = (a < F.n1) && (b < F.n2);
  c.i  = a + F.ngh;
  c.j  = b + F.ngh;
  c.k  = (F.N3 > 1) ? g + F.ngh : 0;
  return c;
}

dim3 active_grid(const FieldView& F) {
  return dim3((F.n1 + BX - 1) / BX, (F.n2 + BY - 1) / BY, F.N3 > 1 ? F.n3 : 1);
}

// ---- Faraday: src/kernels/faraday_mink.hpp:71-166 ---------------------------
struct FD {
  float c1, c2;
  float dx, dy, dz, bxy, byx, bxz, bzx, byz, bzy;
};

template <bool EXT>
__global__ __launch_bounds__(BX* BY) void k_faraday_2d(FieldView F, FD f) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j;
  auto E = [&](int a, int b, int comp) -> float { return F(a, b, 0, comp); };
  if constexpr (!EXT) {
    // standard Yee: alpha = 1, delta = beta = 0
    F(i1, i2, 0, 3) += f.c1 * (-(E(i1, i2 + 1, 2) - E(i1, i2, 2)));
    F(i1, i2, 0, 4) += f.c1 * (+(E(i1 + 1, i2, 2) - E(i1, i2, 2)));
    F(i1, i2, 0, 5) += f.c2 * (+(E(i1, i2 + 1, 0) - E(i1, i2, 0)) - (E(i1 + 1, i2, 1) - E(i1, i2, 1)));
  } else {
    const float alphax = ONE - TWO * f.bxy - THREE * f.dx;
    const float alphay = ONE - TWO * f.byx - THREE * f.dy;
    F(i1, i2, 0, 3) += f.c1 * (-alphay * (E(i1, i2 + 1, 2) - E(i1, i2, 2)) -
                               f.dy * (E(i1, i2 + 2, 2) - E(i1, i2 - 1, 2)) -
                               f.byx * (E(i1 + 1, i2 + 1, 2) - E(i1 + 1, i2, 2)) -
                               f.byx * (E(i1 - 1, i2 + 1, 2) - E(i1 - 1, i2, 2)));
    F(i1, i2, 0, 4) += f.c1 * (+alphax * (E(i1 + 1, i2, 2) - E(i1, i2, 2)) +
                               f.dx * (E(i1 + 2, i2, 2) - E(i1 - 1, i2, 2)) +
                               f.bxy * (E(i1 + 1, i2 + 1, 2) - E(i1, i2 + 1, 2)) +
                               f.bxy * (E(i1 + 1, i2 - 1, 2) - E(i1, i2 - 1, 2)));
    F(i1, i2, 0, 5) += f.c2 * (+alphay * (E(i1, i2 + 1, 0) - E(i1, i2, 0)) +
                               f.dy * (E(i1, i2 + 2, 0) - E(i1, i2 - 1, 0)) +
                               f.byx * (E(i1 + 1, i2 + 1, 0) - E(i1 + 1, i2, 0)) +
                               f.byx * (E(i1 - 1, i2 + 1, 0) - E(i1 - 1, i2, 0)) -
                               alphax * (E(i1 + 1, i2, 1) - E(i1, i2, 1)) -
                               f.dx * (E(i1 + 2, i2, 1) - E(i1 - 1, i2, 1)) -
                               f.bxy * (E(i1 + 1, i2 + 1, 1) - E(i1, i2 + 1, 1)) -
                               f.bxy * (E(i1 + 1, i2 - 1, 1) - E(i1, i2 - 1, 1)));
  }
}

template <bool EXT>
__global__ __launch_bounds__(BX* BY) void k_faraday_3d(FieldView F, FD f) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j, i3 = c.k;
  auto E = [&](int a, int b, int g, int comp) -> float { return F(a, b, g, comp); };
  if constexpr (!EXT) {
    F(i1, i2, i3, 3) += f.c1 * (+(E(i1, i2, i3 + 1, 1) - E(i1, i2, i3, 1)) -
                                (E(i1, i2 + 1, i3, 2) - E(i1, i2, i3, 2)));
    F(i1, i2, i3, 4) += f.c1 * (+(E(i1 + 1, i2, i3, 2) - E(i1, i2, i3, 2)) -
                                (E(i1, i2, i3 + 1, 0) - E(i1, i2, i3, 0)));
    F(i1, i2, i3, 5) += f.c1 * (+(E(i1, i2 + 1, i3, 0) - E(i1, i2, i3, 0)) -
                                (E(i1 + 1, i2, i3, 1) - E(i1, i2, i3, 1)));
  } else {
    const float alphax = ONE - TWO * f.bxy - TWO * f.bxz - THREE * f.dx;
    const float alphay = ONE - TWO * f.byx - TWO * f.byz - THREE * f.dy;
    const float alphaz = ONE - TWO * f.bzx - TWO * f.bzy - THREE * f.dz;
    F(i1, i2, i3, 3) += f.c1 * (
      +alphaz * (E(i1, i2, i3 + 1, 1) - E(i1, i2, i3, 1))
      + f.dz * (E(i1, i2, i3 + 2, 1) - E(i1, i2, i3 - 1, 1))
      + f.bzx * (E(i1 + 1, i2, i3 + 1, 1) - E(i1 + 1, i2, i3, 1))
      + f.bzx * (E(i1 - 1, i2, i3 + 1, 1) - E(i1 - 1, i2, i3, 1))
      + f.bzy * (E(i1, i2 + 1, i3 + 1, 1) - E(i1, i2 + 1, i3, 1))
      + f.bzy * (E(i1, i2 - 1, i3 + 1, 1) - E(i1, i2 - 1, i3, 1))
      - alphay * (E(i1, i2 + 1, i3, 2) - E(i1, i2, i3, 2))
      - f.dy * (E(i1, i2 + 2, i3, 2) - E(i1, i2 - 1, i3, 2))
      - f.byx * (E(i1 + 1, i2 + 1, i3, 2) - E(i1 + 1, i2, i3, 2))
      - f.byx * (E(i1 - 1, i2 + 1, i3, 2) - E(i1 - 1, i2, i3, 2))
      - f.byz * (E(i1, i2 + 1, i3 + 1, 2) - E(i1, i2, i3 + 1, 2))
      - f.byz * (E(i1, i2 + 1, i3 - 1, 2) - E(i1, i2, i3 - 1, 2)));
    F(i1, i2, i3, 4) += f.c1 * (
      +alphax * (E(i1 + 1, i2, i3, 2) - E(i1, i2, i3, 2))
      + f.dx * (E(i1 + 2, i2, i3, 2) - E(i1 - 1, i2, i3, 2))
      + f.bxy * (E(i1 + 1, i2 + 1, i3, 2) - E(i1, i2 + 1, i3, 2))
      + f.bxy * (E(i1 + 1, i2 - 1, i3, 2) - E(i1, i2 - 1, i3, 2))
      + f.bxz * (E(i1 + 1, i2, i3 + 1, 2) - E(i1, i2, i3 + 1, 2))
      + f.bxz * (E(i1 + 1, i2, i3 - 1, 2) - E(i1, i2, i3 - 1, 2))
      - alphaz * (E(i1, i2, i3 + 1, 0) - E(i1, i2, i3, 0))
      - f.dz * (E(i1, i2, i3 + 2, 0) - E(i1, i2, i3 - 1, 0))
      - f.bzx * (E(i1 + 1, i2, i3 + 1, 0) - E(i1 + 1, i2, i3, 0))
      - f.bzx * (E(i1 - 1, i2, i3 + 1, 0) - E(i1 - 1, i2, i3, 0))
      - f.bzy * (E(i1, i2 + 1, i3 + 1, 0) - E(i1, i2 + 1, i3, 0))
      - f.bzy * (E(i1, i2 - 1, i3 + 1, 0) - E(i1, i2 - 1, i3, 0)));
    F(i1, i2, i3, 5) += f.c1 * (
      +alphay * (E(i1, i2 + 1, i3, 0) - E(i1, i2, i3, 0))
      + f.dy * (E(i1, i2 + 2, i3, 0) - E(i1, i2 - 1, i3, 0))
      + f.byx * (E(i1 + 1, i2 + 1, i3, 0) - E(i1 + 1, i2, i3, 0))
      + f.byx * (E(i1 - 1, i2 + 1, i3, 0) - E(i1 - 1, i2, i3, 0))
      + f.byz * (E(i1, i2 + 1, i3 + 1, 0) - E(i1, i2, i3 + 1, 0))
      + f.byz * (E(i1, i2 + 1, i3 - 1, 0) - E(i1, i2, i3 - 1, 0))
      - alphax * (E(i1 + 1, i2, i3, 1) - E(i1, i2, i3, 1))
      - f.dx * (E(i1 + 2, i2, i3, 1) - E(i1 - 1, i2, i3, 1))
      - f.bxy * (E(i1 + 1, i2 + 1, i3, 1) - E(i1, i2 + 1, i3, 1))
      - f.bxy * (E(i1 + 1, i2 - 1, i3, 1) - E(i1, i2 - 1, i3, 1))
      - f.bxz * (E(i1 + 1, i2, i3 + 1, 1) - E(i1, i2, i3 + 1, 1))
      - f.bxz * (E(i1 + 1, i2, i3 - 1, 1) - E(i1, i2, i3 - 1, 1)));
  }
}

// ---- Ampere: src/kernels/ampere_mink.hpp:57-86 ------------------------------
__global__ __launch_bounds__(BX* BY) void k_ampere_2d(FieldView F, float c1, float c2) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j;
  F(i1, i2, 0, 0) += c1 * (F(i1, i2, 0, 5) - F(i1, i2 - 1, 0, 5));
  F(i1, i2, 0, 1) += c1 * (F(i1 - 1, i2, 0, 5) - F(i1, i2, 0, 5));
  F(i1, i2, 0, 2) += c2 * (F(i1, i2 - 1, 0, 3) - F(i1, i2, 0, 3) + F(i1, i2, 0, 4) - F(i1 - 1, i2, 0, 4));
}

__global__ __launch_bounds__(BX* BY) void k_ampere_3d(FieldView F, float c1) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j, i3 = c.k;
  F(i1, i2, i3, 0) += c1 * (F(i1, i2, i3 - 1, 4) - F(i1, i2, i3, 4) + F(i1, i2, i3, 5) -
                            F(i1, i2 - 1, i3, 5));
  F(i1, i2, i3, 1) += c1 * (F(i1 - 1, i2, i3, 5) - F(i1, i2, i3, 5) + F(i1, i2, i3, 3) -
                            F(i1, i2, i3 - 1, 3));
  F(i1, i2, i3, 2) += c1 * (F(i1, i2 - 1, i3, 3) - F(i1, i2, i3, 3) + F(i1, i2, i3, 4) -
                            F(i1 - 1, i2, i3, 4));
}

// ---- E += coeff * J; J /= ppc0: src/kernels/ampere_mink.hpp:156-210 ---------
__global__ __launch_bounds__(BX* BY) void k_currents_ampere(FieldView E, FieldView J, float coeff,
                                                            float ppc0) {
  const CellIdx c = active_cell(E);
  if (!c.ok) return;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const float j = J(c.i, c.j, c.k, comp);
    E(c.i, c.j, c.k, comp) += j * coeff;
    J(c.i, c.j, c.k, comp) = j / ppc0;
  }
}

// ---- digital filter, periodic interior: digital_filter.hpp:182-194, 349-378
__global__ __launch_bounds__(BX* BY) void k_filter_2d(FieldView A, FieldView B) {
  const CellIdx c = active_cell(A);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    auto b = [&](int x, int y) -> float { return B(x, y, 0, comp); };
    A(i1, i2, 0, comp) = INV_4 * b(i1, i2) +
                         INV_8 * (b(i1 - 1, i2) + b(i1 + 1, i2) + b(i1, i2 - 1) + b(i1, i2 + 1)) +
                         INV_16 * (b(i1 - 1, i2 - 1) + b(i1 + 1, i2 + 1) + b(i1 - 1, i2 + 1) +
                                   b(i1 + 1, i2 - 1));
  }
}

// 3D: the 1/32 group is the reference's list as written (it names
// (i,j,k-1) and (i,j,k+1) where the edge pair (i,j-1,k+1), (i,j+1,k-1) would
// complete the binomial stencil).
__global__ __launch_bounds__(BX* BY) void k_filter_3d(FieldView A, FieldView B) {
  const CellIdx c = active_cell(A);
  if (!c.ok) return;
  const int i = c.i, j = c.j, k = c.k;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    auto b = [&](int x, int y, int z) -> float { return B(x, y, z, comp); };
    A(i, j, k, comp) =
      INV_8 * b(i, j, k) +
      INV_16 * (b(i - 1, j, k) + b(i + 1, j, k) + b(i, j - 1, k) + b(i, j + 1, k) + b(i, j, k - 1) +
                b(i, j, k + 1)) +
      INV_32 * (b(i - 1, j - 1, k) + b(i + 1, j + 1, k) + b(i - 1, j + 1, k) + b(i + 1, j - 1, k) +
                b(i, j - 1, k - 1) + b(i, j + 1, k + 1) + b(i, j, k - 1) + b(i, j, k + 1) +
                b(i - 1, j, k - 1) + b(i + 1, j, k + 1) + b(i - 1, j, k + 1) + b(i + 1, j, k - 1)) +
      INV_64 * (b(i - 1, j - 1, k - 1) + b(i + 1, j + 1, k + 1) + b(i - 1, j + 1, k + 1) +
                b(i + 1, j - 1, k - 1) + b(i - 1, j - 1, k + 1) + b(i + 1, j + 1, k - 1) +
                b(i - 1, j + 1, k - 1) + b(i + 1, j - 1, k + 1));
  }
}

// ---- periodic ghost fill: metadomain_comm.cpp:204-367 -----------------------
__device__ __forceinline__ int wrap_src(int i, int G, int n) {
  return i < G ? i + n : (i >= G + n ? i - n : i);
}

__global__ __launch_bounds__(BX* BY) void k_ghost_fill(FieldView F, int c0, int c1) {
  const int i = blockIdx.x * BX + threadIdx.x;
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= F.N1 || j >= F.N2) return;
  const int G  = F.ngh;
  const int si = wrap_src(i, G, F.n1);
  const int sj = wrap_src(j, G, F.n2);
  const int sk = (F.N3 > 1) ? wrap_src(k, G, F.n3) : k;
  if (si == i && sj == j && sk == k) return;
  for (int comp = c0; comp < c1; ++comp) F(i, j, k, comp) = F(si, sj, sk, comp);
}

// ---- periodic J fold: metadomain_comm.cpp:408-562 ---------------------------
// For an active cell, every direction d whose receive slab holds it adds the
// cell's periodic image; images are summed in Directions<D>::all order
// (src/global/arch/directions.h:180-233) and the sum is added to the cell.
__global__ __launch_bounds__(BX* BY) void k_fold(FieldView J) {
  const CellIdx c = active_cell(J);
  if (!c.ok) return;
  const int G = J.ngh;
  const int D = (J.N3 > 1) ? 3 : 2;
  const int n[3] = { J.n1, J.n2, J.n3 };
  const int x[3] = { c.i, c.j, c.k };
  int lo[3], hi[3];  // allowed direction components per axis
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= D) { lo[d] = 0; hi[d] = 0; continue; }
    lo[d] = (x[d] >= n[d]) ? -1 : 0;          // receives from dir -1 : i in [n, n+G)
    hi[d] = (x[d] < 2 * G) ? 1 : 0;           // receives from dir +1 : i in [G, 2G)
  }
  if (lo[0] == 0 && hi[0] == 0 && lo[1] == 0 && hi[1] == 0 && lo[2] == 0 && hi[2] == 0) return;
  for (int comp = 0; comp < 3; ++comp) {
    float acc = ZERO;
    bool  any = false;
    for (int a = lo[0]; a <= hi[0]; ++a)
      for (int b = lo[1]; b <= hi[1]; ++b)
        for (int g = lo[2]; g <= hi[2]; ++g) {
          if (a == 0 && b == 0 && g == 0) continue;
          // direction (a,b,g): source = x + dir * n  (dir +1 -> ghost above)
          acc = any ? acc + J(x[0] + a * n[0], x[1] + b * n[1], x[2] + g * n[2], comp)
                    : ZERO + J(x[0] + a * n[0], x[1] + b * n[1], x[2] + g * n[2], comp);
          any = true;
        }
    J(x[0], x[1], x[2], comp) += acc;
  }
}

ntt_status launch_ok() { return check_launch(); }

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_faraday_mink(ntt_field_t* em, float coeff1, float coeff2,
                                       const ntt_fd_coeffs_t* fd, void* stream) {
  FieldView F;
  ntt_status s = make_field_view(em, 6, &F);
  if (s != NTT_OK) return s;
  FD f{ coeff1, coeff2, 0, 0, 0, 0, 0, 0, 0, 0, 0 };
  bool ext = false;
  if (fd != nullptr) {
    f.dx = fd->delta_x; f.dy = fd->delta_y; f.dz = fd->delta_z;
    f.bxy = fd->beta_xy; f.byx = fd->beta_yx; f.bxz = fd->beta_xz;
    f.bzx = fd->beta_zx; f.byz = fd->beta_yz; f.bzy = fd->beta_zy;
    ext = f.dx != 0 || f.dy != 0 || f.dz != 0 || f.bxy != 0 || f.byx != 0 || f.bxz != 0 ||
          f.bzx != 0 || f.byz != 0 || f.bzy != 0;
  }
  if (ext && F.ngh < 2) return NTT_ERR_ARG;
  const dim3 grid = active_grid(F), block(BX, BY);
  hipStream_t st = as_stream(stream);
  if (em->dim == 2) {
    if (ext) k_faraday_2d<true><<<grid, block, 0, st>>>(F, f);
    else k_faraday_2d<false><<<grid, block, 0, st>>>(F, f);
  } else {
    if (ext) k_faraday_3d<true><<<grid, block, 0, st>>>(F, f);
    else k_faraday_3d<false><<<grid, block, 0, st>>>(F, f);
  }
  return launch_ok();
}

extern "C" ntt_status ntt_ampere_mink(ntt_field_t* em, float coeff1, float coeff2, void* stream) {
  FieldView F;
  ntt_status s = make_field_view(em, 6, &F);
  if (s != NTT_OK) return s;
  const dim3 grid = active_grid(F), block(BX, BY);
  if (em->dim == 2) k_ampere_2d<<<grid, block, 0, as_stream(stream)>>>(F, coeff1, coeff2);
  else k_ampere_3d<<<grid, block, 0, as_stream(stream)>>>(F, coeff1);
  return launch_ok();
}

static bool same_shape(const FieldView& a, const FieldView& b) {
  return a.n1 == b.n1 && a.n2 == b.n2 && a.n3 == b.n3 && a.ngh == b.ngh && a.N3 == b.N3;
}

extern "C" ntt_status ntt_currents_ampere_mink(ntt_field_t* em, ntt_field_t* cur, float coeff,
                                               float ppc0, void* stream) {
  FieldView E, J;
  ntt_status s = make_field_view(em, 6, &E);
  if (s != NTT_OK) return s;
  s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  if (!same_shape(E, J)) return NTT_ERR_ARG;
  k_currents_ampere<<<active_grid(E), dim3(BX, BY), 0, as_stream(stream)>>>(E, J, coeff, ppc0);
  return launch_ok();
}

extern "C" ntt_status ntt_digital_filter(ntt_field_t* array, const ntt_field_t* buffer,
                                         const ntt_flds_bc_t* bc, void* stream) {
  FieldView A, B;
  ntt_status s = make_field_view(array, 3, &A);
  if (s != NTT_OK) return s;
  s = make_field_view(buffer, 3, &B);
  if (s != NTT_OK) return s;
  if (!same_shape(A, B) || A.d == B.d) return NTT_ERR_ARG;
  if (bc != nullptr) {
    // conductor and axis special cases (digital_filter.hpp:129-180) are not built
    for (int d = 0; d < array->dim; ++d)
      for (int e = 0; e < 2; ++e)
        if (bc->bc[d][e] != NTT_FLDS_BC_PERIODIC) return NTT_ERR_UNSUPPORTED;
  }
  if (array->dim == 2) k_filter_2d<<<active_grid(A), dim3(BX, BY), 0, as_stream(stream)>>>(A, B);
  else k_filter_3d<<<active_grid(A), dim3(BX, BY), 0, as_stream(stream)>>>(A, B);
  return launch_ok();
}

extern "C" ntt_status ntt_communicate_fields(ntt_field_t* fld, int32_t comp_begin, int32_t comp_end,
                                             void* stream) {
  FieldView F;
  ntt_status s = make_field_view(fld, 0, &F);
  if (s != NTT_OK) return s;
  if (comp_begin < 0 || comp_end > F.nc || comp_begin >= comp_end) return NTT_ERR_ARG;
  if (F.n1 < F.ngh || F.n2 < F.ngh || (F.N3 > 1 && F.n3 < F.ngh)) return NTT_ERR_ARG;
  const dim3 grid((F.N1 + BX - 1) / BX, (F.N2 + BY - 1) / BY, F.N3);
  k_ghost_fill<<<grid, dim3(BX, BY), 0, as_stream(stream)>>>(F, comp_begin, comp_end);
  return launch_ok();
}

extern "C" ntt_status ntt_synchronize_currents(ntt_field_t* cur, ntt_field_t* buff, void* stream) {
  (void)buff;
  FieldView J;
  ntt_status s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  if (J.n1 < 2 * J.ngh || J.n2 < 2 * J.ngh || (J.N3 > 1 && J.n3 < 2 * J.ngh)) return NTT_ERR_ARG;
  k_fold<<<active_grid(J), dim3(BX, BY), 0, as_stream(stream)>>>(J);
  return launch_ok();
}
