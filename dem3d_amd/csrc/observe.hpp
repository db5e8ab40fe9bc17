// observe.hpp -- bulk observables of a resident state, as deterministic
// reductions on the GPU.
//
//  - KineticEnergy   (src/MDEM.f90:199-213): 1/2 sum (m v.v + I w.w)
//  - PotentialEnergy (src/MDEM.f90:183-195): -sum m (x . Gravity)
//  - the overlapping particle pairs (each once) and particle-wall pairs at
//    the current positions, with the overlap test of the force passes
//    (MinDijSq - DijSq > 0, src/MForces.f90:223-224; Overlap,
//    src/MConfig.f90:692-822), hence the coordination number 2 P / N;
//  - the packing fraction of a bulk region: the volume of the spheres whose
//    centre lies in [zlo, zhi) over Lx Ly (zhi - zlo), the region being the
//    whole box when z is periodic, else the slab from one largest diameter
//    above the bottom of the box to one largest diameter below the top of the
//    packing (max z + r), which leaves out the layers a flat wall orders and
//    the thinned free surface;
//  - optionally a vertical profile over nl equal layers of the box height:
//    the volume of sphere material inside each layer (exact sphere-slab
//    intersection) over the layer volume, and the volume-weighted mean
//    velocity of that material.
//
// The reference driver does not compute the last three; they are defined
// here for the statements about packings and the mixed-precision mode.
//
// Every sum has a fixed order: each block reduces its particles with a fixed
// tree, blocks stride over the particles in a fixed pattern, and one block
// adds the block partials in index order.  The results are bitwise the same
// from run to run and whatever memory holds the partials.
//
// Scratch layout (doubles): block partials [DEM_OBS_BLOCKS][DEM_OBS_NPART],
// layer partials [DEM_OBS_BLOCKS][nl][4], results [DEM_OBS_NOUT + 4 nl].
#pragma once

#include "kernels.hpp"

#define DEM_OBS_BLOCKS 256
#define DEM_OBS_NPART 8
#define DEM_OBS_NOUT 16
#define DEM_OBS_PI 3.1415926535897932384626433832795

// result slots
enum {
  OBS_KE = 0,
  OBS_PE = 1,
  OBS_TOP = 2,
  OBS_PAIRS = 3,
  OBS_WALLS = 4,
  OBS_ZLO = 5,
  OBS_ZHI = 6,
  OBS_VOL = 7,
  OBS_PHI = 8,
};

struct ObsArgs {
  const double4* pos;
  const Kin* kin;
  EdgeSets E;
  int nl;           // layers of the profile (0: none)
  int nblk;         // blocks of the partial-sum kernels
  double* part;     // [DEM_OBS_BLOCKS][DEM_OBS_NPART]
  double* lpart;    // [DEM_OBS_BLOCKS][nl][4]
  double* out;      // [DEM_OBS_NOUT + 4 nl]
};

__host__ __device__ __forceinline__ size_t obs_scratch_doubles(int nl) {
  return (size_t)DEM_OBS_BLOCKS * DEM_OBS_NPART + (size_t)DEM_OBS_BLOCKS * (size_t)nl * 4 +
         DEM_OBS_NOUT + 4 * (size_t)nl;
}

// Volume of the sphere (centre z, radius r) between heights a < b.
__device__ __forceinline__ double slab_volume(double z, double r, double a, double b) {
  const double lo = fmax(a, z - r), hi = fmin(b, z + r);
  if (!(hi > lo)) return 0.0;
  const double u = lo - z, v = hi - z;
  return DEM_OBS_PI * (r * r * (v - u) - (v * v * v - u * u * u) / 3.0);
}

__device__ __forceinline__ double block_max_det(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int t = threadIdx.x;
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double r = -INFINITY;
  if (t == 0)
    for (int w = 0; w < DEM_BLOCK / 64; ++w) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}

// Pass 1: per-block partials of the energies, the top of the packing, the
// contact counts and the layer sums.  Block b handles the 256-particle
// chunks b, b + nblk, b + 2 nblk, ...
__global__ __launch_bounds__(DEM_BLOCK) void k_obs_partial(SysConst s, ObsArgs o) {
  __shared__ double sh[DEM_BLOCK / 64];
  __shared__ int kr[2][DEM_BLOCK / 64];
  const int t = threadIdx.x, b = blockIdx.x;
  const int nl = o.nl;
  const double lz = s.lbox[2], zb = -0.5 * lz, dz = nl > 0 ? lz / (double)nl : 1.0;
  double* lrow = o.lpart + (size_t)b * (size_t)nl * 4;
  for (int q = t; q < 4 * nl; q += DEM_BLOCK) lrow[q] = 0.0;
  __syncthreads();
  const int ep = *o.E.epar;
  double ke = 0.0, pe = 0.0, top = -INFINITY, npair = 0.0, nwall = 0.0;
  const int nchunk = (s.np + DEM_BLOCK - 1) / DEM_BLOCK;
  for (int c = b; c < nchunk; c += o.nblk) {
    const int i = c * DEM_BLOCK + t;
    const bool active = i < s.np;
    double kei = 0.0, pei = 0.0, topi = -INFINITY, zi = 0.0, ri = 0.0;
    d3 vi = mk3(0.0, 0.0, 0.0);
    unsigned np_i = 0, nw_i = 0;
    int klo = 0x7fffffff, khi = -1;
    if (active) {
      const double4 P = o.pos[i];
      const Kin K = load_kin(o.kin + i);
      const int ci = (int)P.w;
      const CompConst cc = s.comp[ci];
      vi = kin_v(K);
      const d3 w = kin_w(K);
      // KineticEnergy term (src/MDEM.f90:209-210)
      kei = cc.mass * dot3(vi, vi) + cc.inertia * dot3(w, w);
      // PotentialEnergy term (src/MDEM.f90:191-192)
      const d3 x = mk3(P.x, P.y, P.z);
      pei = cc.mass * dot3(x, mk3(s.grav[0], s.grav[1], s.grav[2]));
      ri = cc.radius;
      zi = P.z;
      topi = P.z + ri;
      // directed overlapping edges (each pair appears in both lists)
      const int cn = o.E.cnt[ep][i];
      const long long e = sell_edge(o.E.soff[ep], i, 0);
      for (int k = 0; k < cnt_deg(cn); ++k) {
        const double4 Q = o.pos[o.E.nbr[ep][e + 64LL * k]];
        const d3 r = min_image3(s, mk3(Q.x - P.x, Q.y - P.y, Q.z - P.z));
        if (s.pair[ci + (int)Q.w * s.nc].mindijsq - dot3(r, r) > 0.0) ++np_i;
      }
      for (int jw = 0; jw < s.nw; ++jw) {
        const WallConst& wc = s.wallc[ci + jw * s.nc];
        const WallHit wh = wall_overlap(s, s.walls[jw], x, ri, wc.max_alpha_u, wc.min_aw_sq,
                                        wc.max_aw_sq);
        if (wh.ov > 0) ++nw_i;
      }
      if (nl > 0) {
        // layers the sphere cuts, clamped to the box (and finite for any z)
        klo = (int)fmax(0.0, fmin(floor((zi - ri - zb) / dz), (double)(nl - 1)));
        khi = (int)fmax(0.0, fmin(floor((zi + ri - zb) / dz), (double)(nl - 1)));
      }
    }
    ke += block_sum_det(kei, sh);
    pe += block_sum_det(pei, sh);
    top = fmax(top, block_max_det(topi, sh));
    npair += block_sum_det((double)np_i, sh);
    nwall += block_sum_det((double)nw_i, sh);
    if (nl > 0) {
      // layers this chunk touches
      for (int q = 32; q > 0; q >>= 1) {
        klo = min(klo, __shfl_xor(klo, q, 64));
        khi = max(khi, __shfl_xor(khi, q, 64));
      }
      if ((t & 63) == 0) {
        kr[0][t >> 6] = klo;
        kr[1][t >> 6] = khi;
      }
      __syncthreads();
      int bl = kr[0][0], bh = kr[1][0];
      for (int w = 1; w < DEM_BLOCK / 64; ++w) {
        bl = min(bl, kr[0][w]);
        bh = max(bh, kr[1][w]);
      }
      __syncthreads();
      for (int k = bl; k <= bh; ++k) {
        const double v = active ? slab_volume(zi, ri, zb + (double)k * dz,
                                              zb + (double)(k + 1) * dz)
                                : 0.0;
        const double s0 = block_sum_det(v, sh);
        const double s1 = block_sum_det(v * vi.x, sh);
        const double s2 = block_sum_det(v * vi.y, sh);
        const double s3 = block_sum_det(v * vi.z, sh);
        if (t == 0) {
          double* q = lrow + 4 * (size_t)k;
          q[0] += s0;
          q[1] += s1;
          q[2] += s2;
          q[3] += s3;
        }
      }
    }
  }
  if (t == 0) {
    double* q = o.part + (size_t)b * DEM_OBS_NPART;
    q[0] = ke;
    q[1] = pe;
    q[2] = top;
    q[3] = npair;
    q[4] = nwall;
  }
}

// Pass 2 (one block): the totals, the bulk region and the profile.
__global__ __launch_bounds__(DEM_BLOCK) void k_obs_finish(SysConst s, ObsArgs o) {
  __shared__ double sh[DEM_BLOCK / 64];
  const int t = threadIdx.x;
  const bool has = t < o.nblk;
  const double* q = o.part + (size_t)t * DEM_OBS_NPART;
  const double ke = block_sum_det(has ? q[0] : 0.0, sh);
  const double pe = block_sum_det(has ? q[1] : 0.0, sh);
  const double top = block_max_det(has ? q[2] : -INFINITY, sh);
  const double npair = block_sum_det(has ? q[3] : 0.0, sh);
  const double nwall = block_sum_det(has ? q[4] : 0.0, sh);
  if (t == 0) {
    double dmax = 0.0;
    for (int c = 0; c < s.nc; ++c) dmax = fmax(dmax, 2.0 * s.comp[c].radius);
    double zlo = -0.5 * s.lbox[2], zhi = 0.5 * s.lbox[2];
    if (!s.pbc[2] && s.np > 0) {
      const double a = zlo + dmax, b = top - dmax;
      if (b > a) {
        zlo = a;
        zhi = b;
      }
    }
    o.out[OBS_KE] = 0.5 * ke;
    o.out[OBS_PE] = -pe;
    o.out[OBS_TOP] = top;
    o.out[OBS_PAIRS] = 0.5 * npair;
    o.out[OBS_WALLS] = nwall;
    o.out[OBS_ZLO] = zlo;
    o.out[OBS_ZHI] = zhi;
  }
  const int nl = o.nl;
  const double dz = nl > 0 ? s.lbox[2] / (double)nl : 1.0;
  const double lvol = s.lbox[0] * s.lbox[1] * dz;
  for (int k = t; k < nl; k += DEM_BLOCK) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int b = 0; b < o.nblk; ++b) {
      const double* r = o.lpart + ((size_t)b * (size_t)nl + (size_t)k) * 4;
      a0 += r[0];
      a1 += r[1];
      a2 += r[2];
      a3 += r[3];
    }
    double* p = o.out + DEM_OBS_NOUT + 4 * (size_t)k;
    p[0] = a0 / lvol;
    p[1] = a0 > 0.0 ? a1 / a0 : 0.0;
    p[2] = a0 > 0.0 ? a2 / a0 : 0.0;
    p[3] = a0 > 0.0 ? a3 / a0 : 0.0;
  }
}

// Pass 3: per-block volumes of the spheres whose centre lies in the region.
__global__ __launch_bounds__(DEM_BLOCK) void k_obs_bulk(SysConst s, ObsArgs o) {
  __shared__ double sh[DEM_BLOCK / 64];
  const int t = threadIdx.x, b = blockIdx.x;
  const double zlo = o.out[OBS_ZLO], zhi = o.out[OBS_ZHI];
  double vol = 0.0;
  const int nchunk = (s.np + DEM_BLOCK - 1) / DEM_BLOCK;
  for (int c = b; c < nchunk; c += o.nblk) {
    const int i = c * DEM_BLOCK + t;
    double v = 0.0;
    if (i < s.np) {
      const double4 P = o.pos[i];
      if (P.z >= zlo && P.z < zhi) {
        const double r = s.comp[(int)P.w].radius;
        v = 4.0 / 3.0 * DEM_OBS_PI * ((r * r) * r);
      }
    }
    vol += block_sum_det(v, sh);
  }
  if (t == 0) o.part[(size_t)b * DEM_OBS_NPART + 5] = vol;
}

// Pass 4 (one block): the packing fraction.
__global__ __launch_bounds__(DEM_BLOCK) void k_obs_phi(SysConst s, ObsArgs o) {
  __shared__ double sh[DEM_BLOCK / 64];
  const int t = threadIdx.x;
  const double vol =
      block_sum_det(t < o.nblk ? o.part[(size_t)t * DEM_OBS_NPART + 5] : 0.0, sh);
  if (t == 0) {
    o.out[OBS_VOL] = vol;
    const double h = o.out[OBS_ZHI] - o.out[OBS_ZLO];
    o.out[OBS_PHI] = vol / (s.lbox[0] * s.lbox[1] * h);
  }
}
