// Integer-split (Ozaki-type) int8 MFMA kernels (see ozaki.hip).
#pragma once
#include "common.hpp"

#include <vector>

namespace nopt {
// Rows of src (rows_src x cols, ld) -> ns digit planes
// planes[k*pstride + row*ldo + col] (rows_dst >= rows_src rows, extra rows and
// columns zero) and one exponent per row.
// Optional reorders: output row r reads source row (r/rper)*rper +
// rperm[r%rper], output column c source column cperm[c].
hipError_t split_rows_i8(const double* src, long ld, int cols, long rows_src, int ns, int8_t* planes, long ldo,
                         long rows_dst, long pstride, int* expo, hipStream_t s, const int* rperm = nullptr,
                         long rper = 0, const int* cperm = nullptr);
// Full rows from a chunk of a symmetric tensor kept as packed lower
// triangles, A[P*cp + m(m+1)/2 + n - m0(m0+1)/2] for the rows m0 <= m < m1
// (n <= m) of every P < naux: each element goes to both full rows it belongs
// to, (P, m) and (P, n). unpack_packed_f64 writes FP64 rows B[(P*nbf + r)*ld + col];
// rowmax_packed folds the chunk into the row maxima (bit patterns, zeroed by
// the caller), rowmax_exponents turns them into exponents, and
// split_packed_i8 writes the digit planes of split_rows_i8 with those
// exponents. Padding columns are the caller's to clear. With pinv, the
// function m of the chunk is function pinv[m] of the rows written (rows and
// columns both reordered).
hipError_t unpack_packed_f64(const double* A, long cp, int m0, int m1, int nbf, long naux, double* B, long ld,
                             hipStream_t s);
hipError_t rowmax_packed(const double* A, long cp, int m0, int m1, int nbf, long naux, unsigned long long* rowmax,
                         hipStream_t s, const int* pinv = nullptr);
hipError_t rowmax_exponents(const unsigned long long* rowmax, long rows, int* expo, hipStream_t s);
hipError_t split_packed_i8(const double* A, long cp, int m0, int m1, int nbf, long naux, const int* expo, int ns,
                           int8_t* planes, long ldo, long pstride, hipStream_t s, const int* pinv = nullptr);
// Digit split of the half-transformed rows fused into the half-transform:
// segment d covers the columns [off[d], off[d] + len[d]) of X (off a multiple
// of 16); its rows get their own exponent xe[d][row] and four digit planes
// xd[d][p*xplane[d] + row*ldk[d] + col], zero beyond the written columns
// (ldk[d] a multiple of 64; rows and columns as launch_half_i8 states).
constexpr int kMaxXSeg = 8;
struct XSplit {
  int nseg;
  int off[kMaxXSeg], len[kMaxXSeg];
  long ldk[kMaxXSeg], xplane[kMaxXSeg];
  int8_t* xd[kMaxXSeg];
  int* xe[kMaxXSeg];
};

// B operand of the half-transform: the full rows as four digit planes
// (Bd[k*bplane + (Q*nbf + m)*ld8 + n]), or, with Bd null, the packed lower
// triangle of the Coulomb form below (Bp = (const int*)Bs, its N28 values,
// sq values per function) with the row exponents also as bytes Be8[Q*ld8 + n].
// Be holds the row exponents in either case.
struct HalfB {
  const int8_t* Bd;
  long bplane;
  const int* Bp;
  long sq;
  const int8_t* Be8;
  const int* Be;
  // optional: Bm[Q*nmb + mb] bit k set when the rows 32 mb .. 32 mb + 31 of
  // function Q hold a nonzero digit in the columns 64 k .. 64 k + 63, as the
  // kernel reads them (slab_mask_rows / slab_mask_packed); null: every slab
  const uint32_t* Bm = nullptr;
  int nmb = 0;
};

// Half-transform X = B C^T (FP64 out) from digit planes of B and C
// (grid: row tiles (the launcher picks their height) x ceil(naux/qpb) x
// column groups of 16*nt columns, nt in {2,4,6}; ld8 a multiple of 64, at
// most 2048 (nbf <= 2048: the int32 class combine stays exact); the
// caller's grid.x is ignored).
// With xs.nseg > 0 the rows are split into digit planes in the epilogue and
// X is not written: segment d covers the columns [off[d], off[d] + len[d])
// (global column numbers, within one column group). qpb = 2 gives each block
// the functions 2y and 2y+1, whose rows (y*nbf + m) share one exponent, the
// function 2y in columns [0, len) and 2y+1 in [len, 2 len) of the segment
// (ldk >= 2 len); qpb = 1 writes row Q*nbf + m.
hipError_t launch_half_i8(int nt, int qpb, dim3 grid, hipStream_t s, const HalfB& b, long ld8, int nbf, int naux,
                          const int8_t* Cd, long cplane, const int* Ce, double* X, long ldx, const XSplit& xs);
// Coulomb pair over B's lower triangle in packed rows: row m of B_Q as
// m/16 + 1 segments of 16 columns (zeros beyond column m), the rows one after
// the other, element e = Q*sq + 16*seg0(m) + n, sq = 16*nseg; each value as
// its 35-bit integer N35 = 128 N28 + d4, ((int*)Bs)[e] = N28 and
// ((int8_t*)Bs)[4*splane + e] = d4 (pack_segs_rows forms them from FP64
// rows, pack_segs_chunk from a packed chunk of rows m0..m1; seg_code[s] =
// m << 16 | n0 for each segment; Bs holds 5*splane bytes). expo_bytes writes
// the row exponents as bytes. The kernels walk 8-column slots (m, n0)
// of the triangle (slot table, m/8 + 1 slots per row): gamma (partials per
// slot block, then summed) and J partials per Q-split
// (Jpart[d*jstride + split*nbf*nbf + m*nbf + n], n <= m).
constexpr int kGammaSPL = 2;  // slots per lane of the gamma kernel
constexpr int kJSPL = 2;      // slots per thread of the J kernel
int coulomb_seg_count(int nbf);
void coulomb_seg_table(int nbf, int* out);
hipError_t pack_segs_rows(const double* B, long ld, int nbf, long naux, const int* Be, const int* seg_code, int nseg,
                          int8_t* Bs, long splane, hipStream_t s, const int* perm = nullptr);
hipError_t pack_segs_chunk(const double* A, long cp, int m0, int m1, int nbf, long naux, const int* Be,
                           const int* seg_code, int nseg, int8_t* Bs, long splane, hipStream_t s);
// pack_segs_rows with perm: Bs of the reordered functions (row and column m
// of Bs read row and column perm[m] of B); scatter_segs_chunk: the same from
// a packed chunk, each stored element to its place in the reordered
// triangle (pinv = the inverse order; Bs zeroed before the first chunk).
hipError_t scatter_segs_chunk(const double* A, long cp, int m0, int m1, int nbf, long naux, const int* Be,
                              const int* pinv, int nseg, int8_t* Bs, long splane, hipStream_t s);
hipError_t expo_bytes(const int* Be, int nbf, long ld8, long naux, int8_t* Be8, hipStream_t s);
int coulomb_slot_count(int nbf);
void coulomb_slot_table(int nbf, int* out);
// Blocks of exact zeros (values below their row's resolution are stored as
// zeros). seg_live: live[s] |= 1 (2) when the first (last) eight columns of
// segment s hold a nonzero N28 or d4 for some Q (live zeroed before); the
// Coulomb kernels then walk tables of the live segments and slots only.
// slab_mask_rows / slab_mask_packed: the half-transform's slab masks (HalfB::Bm)
// from the full-row planes or from the N28 of Bs, nmb = ceil(nbf/32) words
// per function.
hipError_t seg_live(const int8_t* Bs, long splane, int nseg, long naux, int* live, hipStream_t s);
hipError_t slab_mask_rows(const int8_t* Bd, long bplane, long ld8, int nbf, long naux, uint32_t* Bm, hipStream_t s);
hipError_t slab_mask_packed(const int* Bp, long sq, int nbf, long naux, uint32_t* Bm, hipStream_t s);
int gamma_slot_blocks(int nslots);
// One or two densities walk the 16-column segments instead (one 16-byte load
// per plane and lane); gamma_partial_blocks bounds the partials of either.
hipError_t launch_gamma_slots(int nd, int nsplit, hipStream_t s, const int8_t* Bs, long splane, long sq, const int* Be,
                              int nbf, int naux, const int* slots, int nslots, const int* segs, int nseg,
                              const double* Dp, long ldd, long dstride, double* gpart, double* gamma, long gstride);
int gamma_partial_blocks(int nslots, int nseg);
hipError_t launch_jpart_slots(int nd, int nsplit, hipStream_t s, const int8_t* Bs, long splane, long sq, const int* Be,
                              int nbf, int naux, const int* slots, int nslots, const int* segs, int nseg,
                              const double* gamma, long gstride, double* Jpart, long jstride);

// Exchange partials over 64x64 lower-triangle tiles: Kpart[split][tile] for
// ntiles tiles and nsplit Q-splits (each XCD a contiguous range of the
// (tile, split) pairs).
constexpr int kSyrkI8Tile = 64;
hipError_t launch_syrk_i8(int ntiles, int nsplit, hipStream_t s, const int8_t* Xd, long xplane, const int* Xe,
                          long ldk, int nbf, int naux, double* Kpart);

// Work lists (common.hpp): the kernels above can run the grids of several
// systems in one launch. A ListRec is one item: the kernel (kind and template
// configuration), its arguments and its number of blocks. The *_item
// functions check their arguments as the launchers above do and fill recs
// instead of launching; launch_list runs n items of one kind and
// configuration, given their arguments (an array of the kind's struct) and
// first[i], the first block of item i (first[n] = blocks).
struct ListRec {
  int kind = 0, cfg = 0, blocks = 0, threads = 0;
  std::vector<unsigned char> args;
};
enum ListKind { kListSplit = 0, kListHalf, kListGamma, kListGammaSum, kListJ, kListSyrk, kListOzakiKinds };
struct ItemSplit {
  const double* src;
  long ld;
  int cols;
  long rows_src;
  int ns;
  int8_t* planes;
  long ldo;
  long pstride;
  int* expo;
  const int* cperm;
  int gx, gy, gz;
};
struct ItemHalf {
  const int8_t* Bd;
  long bplane;
  const int* Bp;
  long sq;
  const int8_t* Be8;
  const int* Be;
  const uint32_t* Bm;
  int nmb;
  long ld8;
  int nbf;
  int naux;
  const int8_t* Cd;
  long cplane;
  const int* Ce;
  double* X;
  long ldx;
  XSplit xs;
  int gx, gy, gz;
};
struct ItemGamma {
  const int8_t* Bs;
  long splane;
  long sq;
  const int* Be;
  int nbf;
  int naux;
  const int* slots;
  int nslots;
  const double* Dp;
  long ldd;
  long dstride;
  double* gpart;
  int gx, gy, gz;
};
struct ItemGammaSum {
  const double* gpart;
  int nsb;
  int nd;
  int naux;
  double* gamma;
  long gstride;
  int gx, gy, gz;
};
struct ItemJ {
  const int8_t* Bs;
  long splane;
  long sq;
  const int* Be;
  int nbf;
  int naux;
  const int* slots;
  int nslots;
  const double* gamma;
  long gstride;
  double* Jpart;
  long jstride;
  int gx, gy, gz;
};
struct ItemSyrk {
  const int8_t* Xd;
  long xplane;
  const int* Xe;
  long ldk;
  int nbf;
  int naux;
  int ntiles;
  int nsplit;
  double* Kpart;
  int gx, gy, gz;
};
hipError_t split_rows_item(const double* src, long ld, int cols, long rows_src, int ns, int8_t* planes, long ldo,
                           long rows_dst, long pstride, int* expo, ListRec* r, const int* cperm = nullptr);
hipError_t half_i8_item(int nt, int qpb, dim3 grid, const HalfB& b, long ld8, int nbf, int naux, const int8_t* Cd,
                        long cplane, const int* Ce, double* X, long ldx, const XSplit& xs, ListRec* r);
// two recs: the gamma partials and their sum
hipError_t gamma_slots_items(int nd, int nsplit, const int8_t* Bs, long splane, long sq, const int* Be, int nbf,
                             int naux, const int* slots, int nslots, const int* segs, int nseg, const double* Dp,
                             long ldd, long dstride, double* gpart, double* gamma, long gstride, ListRec* r);
hipError_t jpart_slots_item(int nd, int nsplit, const int8_t* Bs, long splane, long sq, const int* Be, int nbf,
                            int naux, const int* slots, int nslots, const int* segs, int nseg, const double* gamma,
                            long gstride, double* Jpart, long jstride, ListRec* r);
hipError_t syrk_i8_item(int ntiles, int nsplit, const int8_t* Xd, long xplane, const int* Xe, long ldk, int nbf,
                        int naux, double* Kpart, ListRec* r);
hipError_t launch_list(int kind, int cfg, const void* items, const int* first, int n, int blocks, hipStream_t s);
}  // namespace nopt
