// hgx_k_row_tr.hip -- row-pass kernels of the transform-only entry points (ntt-forward / ntt-reverse).
#include "hgx_kernels.hpp"
#include "hgx_launch.hpp"

bool hgx_launch_row_tr(int mode, int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                       const JobRef &jr, int jt_bits) {
  if (mode == MODE_FWD) {
    switch (sbits) {
    case 1: hipLaunchKernelGGL((k_row<1, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 2: hipLaunchKernelGGL((k_row<2, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 3: hipLaunchKernelGGL((k_row<3, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 4: hipLaunchKernelGGL((k_row<4, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 5: hipLaunchKernelGGL((k_row<5, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 6: hipLaunchKernelGGL((k_row<6, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 7: hipLaunchKernelGGL((k_row<7, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 8: hipLaunchKernelGGL((k_row<8, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 9: hipLaunchKernelGGL((k_row<9, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 10: hipLaunchKernelGGL((k_row<10, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 11: hipLaunchKernelGGL((k_row<11, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 12: hipLaunchKernelGGL((k_row<12, MODE_FWD, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    }
  } else {
    switch (sbits) {
    case 1: hipLaunchKernelGGL((k_row<1, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 2: hipLaunchKernelGGL((k_row<2, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 3: hipLaunchKernelGGL((k_row<3, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 4: hipLaunchKernelGGL((k_row<4, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 5: hipLaunchKernelGGL((k_row<5, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 6: hipLaunchKernelGGL((k_row<6, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 7: hipLaunchKernelGGL((k_row<7, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 8: hipLaunchKernelGGL((k_row<8, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 9: hipLaunchKernelGGL((k_row<9, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 10: hipLaunchKernelGGL((k_row<10, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 11: hipLaunchKernelGGL((k_row<11, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    case 12: hipLaunchKernelGGL((k_row<12, MODE_INV, false>), grid, dim3(HGX_NT), 0, s, ps, rp, jr, jt_bits); return true;
    }
  }
  return false;
}

bool hgx_launch_row_inv_scaled(int sbits, dim3 grid, hipStream_t s, const PrimeSet &ps, const RowParams &rp,
                               const JobRef &jr) {
  switch (sbits) {
    case 1: hipLaunchKernelGGL((k_row_inv_scaled<1>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 2: hipLaunchKernelGGL((k_row_inv_scaled<2>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 3: hipLaunchKernelGGL((k_row_inv_scaled<3>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 4: hipLaunchKernelGGL((k_row_inv_scaled<4>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 5: hipLaunchKernelGGL((k_row_inv_scaled<5>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 6: hipLaunchKernelGGL((k_row_inv_scaled<6>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 7: hipLaunchKernelGGL((k_row_inv_scaled<7>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 8: hipLaunchKernelGGL((k_row_inv_scaled<8>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 9: hipLaunchKernelGGL((k_row_inv_scaled<9>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 10: hipLaunchKernelGGL((k_row_inv_scaled<10>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 11: hipLaunchKernelGGL((k_row_inv_scaled<11>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
    case 12: hipLaunchKernelGGL((k_row_inv_scaled<12>), grid, dim3(HGX_NT), 0, s, ps, rp, jr); return true;
  }
  return false;
}
