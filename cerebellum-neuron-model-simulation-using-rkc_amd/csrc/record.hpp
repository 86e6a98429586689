// Device-side output of one population after each output interval (the
// reference copies every compartment's V to the host after every macro step,
// main.cu:263-273, and finds the raster crossings there, pkj.cu:370-412):
// the recorded voltages and the spike bits of an interval are written on the
// device and copied back once per run.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cbrkc {

// One thread per cell: rec[c] = V of compartment rec_comp (rec_comp >= 0), or
// rec[c * comp + q] = V of every compartment q (rec_comp < 0); with bits, an
// upward crossing of thr by compartment raster_comp since the previous
// interval sets bit c % 32 of bits[c / 32] (prev starts at 100 mV, as the
// reference's raster bookkeeping, main.cu:162-169).
static __global__ void __launch_bounds__(128) k_rec_step(const double *__restrict__ v, int comp, int n, int rec_comp,
                                                  double *__restrict__ rec, int raster_comp, double thr,
                                                  double *__restrict__ prev, uint32_t *__restrict__ bits) {
  const int c = blockIdx.x * 128 + threadIdx.x;
  if (c >= n) return;
  const double *vc = v + (size_t)c * comp;
  if (rec) {
    if (rec_comp >= 0) {
      rec[c] = vc[rec_comp];
    } else {
      for (int q = 0; q < comp; q++) rec[(size_t)c * comp + q] = vc[q];
    }
  }
  if (bits) {
    const double x = vc[raster_comp];
    if (x > thr && prev[c] <= thr) atomicOr(&bits[c >> 5], 1u << (c & 31));
    prev[c] = x;
  }
}

}  // namespace cbrkc
