// hgx_k_row_sqr.hip -- row-pass kernels of squares (one operand, pointwise square).
#include "hgx_kernels.hpp"
#include "hgx_launch.hpp"

bool hgx_launch_row_sqr(bool single, int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                        const JobRef &jr, int jt_bits) {
  if (single) {
    switch (sbits) {
    case 1: hipLaunchKernelGGL((k_row<1, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 2: hipLaunchKernelGGL((k_row<2, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 3: hipLaunchKernelGGL((k_row<3, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 4: hipLaunchKernelGGL((k_row<4, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 5: hipLaunchKernelGGL((k_row<5, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 6: hipLaunchKernelGGL((k_row<6, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 7: hipLaunchKernelGGL((k_row<7, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 8: hipLaunchKernelGGL((k_row<8, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 9: hipLaunchKernelGGL((k_row<9, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 10: hipLaunchKernelGGL((k_row<10, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 11: hipLaunchKernelGGL((k_row<11, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 12: hipLaunchKernelGGL((k_row<12, MODE_SQR, true>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    }
  } else {
    switch (sbits) {
    case 1: hipLaunchKernelGGL((k_row<1, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 2: hipLaunchKernelGGL((k_row<2, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 3: hipLaunchKernelGGL((k_row<3, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 4: hipLaunchKernelGGL((k_row<4, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 5: hipLaunchKernelGGL((k_row<5, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 6: hipLaunchKernelGGL((k_row<6, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 7: hipLaunchKernelGGL((k_row<7, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 8: hipLaunchKernelGGL((k_row<8, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 9: hipLaunchKernelGGL((k_row<9, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 10: hipLaunchKernelGGL((k_row<10, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 11: hipLaunchKernelGGL((k_row<11, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 12: hipLaunchKernelGGL((k_row<12, MODE_SQR, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    }
  }
  return false;
}

// small products of one length, several per tile (k_row_pack)
bool hgx_launch_row_pack_sqr(int sbits, int jbits, unsigned njobs, dim3 grid, hipStream_t s, const PrimeSet &ps,
                           const RowParams &rp, const JobRef &jr) {
  if (jbits < 0 || sbits + jbits > 12) return false;
  switch (sbits) {
  case 4: hipLaunchKernelGGL((k_row_pack<4, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  case 5: hipLaunchKernelGGL((k_row_pack<5, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  case 6: hipLaunchKernelGGL((k_row_pack<6, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  case 7: hipLaunchKernelGGL((k_row_pack<7, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  case 8: hipLaunchKernelGGL((k_row_pack<8, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  case 9: hipLaunchKernelGGL((k_row_pack<9, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  case 10: hipLaunchKernelGGL((k_row_pack<10, MODE_SQR>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jbits, njobs); return true;
  }
  return false;
}
