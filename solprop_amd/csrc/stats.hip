// Plane statistics (see stats.hpp): two reductions, both in a fixed order.
#include "stats.hpp"

#include <cmath>
#include <limits>

#include "kernels.hpp"

namespace sp {

namespace {

constexpr int kStatThreads = 256;
constexpr int kSums = 3 + kMaxStatWindows;  // weighted, plain and omega-weighted |X|^2

// One workgroup per column: sum_k w_k |X_k|^2 (the energy's weights: dt / n_t, times 4
// on a field grid's bins strictly between DC and Nyquist), sum_k |X_k|^2 and
// sum_k omega_k |X_k|^2 (the centroid's, whose density prefactor cancels), and the
// weighted sum inside each half-open window lo <= omega < hi. Each thread strides the
// bins in order; the block then reduces in a fixed tree.
__global__ void __launch_bounds__(kStatThreads)
    stat_spectral_sums(const cplx* y, int N, int n_t, int field, double dt,
                       const double* omega, const double* windows, int n_windows,
                       double* sums) {
  __shared__ double red[kStatThreads];
  const int c = blockIdx.x;
  const cplx* row = y + static_cast<long>(c) * N;
  double acc[kSums];
  for (int j = 0; j < kSums; ++j) acc[j] = 0.0;
  const double unit = dt / n_t;
  for (int k = threadIdx.x; k < N; k += kStatThreads) {
    const double p = row[k].x * row[k].x + row[k].y * row[k].y;
    const bool edge = k == 0 || (field && 2 * (N - 1) == n_t && k == N - 1);
    const double w = (field && !edge) ? 4.0 * unit : unit;
    const double wp = w * p;
    acc[0] += wp;
    acc[1] += p;
    acc[2] += omega[k] * p;
    // Unrolled over the largest window count: the accumulators stay in registers.
#pragma unroll
    for (int j = 0; j < kMaxStatWindows; ++j)
      if (j < n_windows && omega[k] >= windows[2 * j] && omega[k] < windows[2 * j + 1])
        acc[3 + j] += wp;
  }
#pragma unroll
  for (int j = 0; j < kSums; ++j) {
    if (j < 3 + n_windows) {  // uniform across the workgroup
      red[threadIdx.x] = acc[j];
      __syncthreads();
      for (int width = kStatThreads / 2; width > 0; width >>= 1) {
        if (threadIdx.x < width) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + width];
        __syncthreads();
      }
      if (threadIdx.x == 0) sums[static_cast<long>(c) * kSums + j] = red[0];
      __syncthreads();
    }
  }
}

// The analytic signal's spectrum on the coarse grid, 1/n_t folded in for the unscaled
// inverse transform: a field grid's one-sided bins doubled except DC and Nyquist, the
// negative frequencies zero (src/solprop/fourier.py:243-300); an envelope's bins as
// they are (src/solprop/stats/definitions.py:148-178).
// With `gate` (lo, hi) given, the spectrum is first gated to lo <= omega < hi, as
// peak_power_window gates it (definitions.py:725-760).
__global__ void stat_analytic_pack(const cplx* y, int C, int N, int n_t, int field,
                                   const double* omega, const double* gate, cplx* out) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(C) * n_t) return;
  const int c = static_cast<int>(idx / n_t), k = static_cast<int>(idx % n_t);
  const double inv = 1.0 / n_t;
  cplx v = make(0.0, 0.0);
  const bool open = k < N && (!gate || (omega[k] >= gate[0] && omega[k] < gate[1]));
  if (!field) {
    if (open) v = scale(inv, y[static_cast<long>(c) * N + k]);
  } else if (k <= n_t / 2 && open) {
    const double h = (k == 0 || k == n_t / 2) ? 1.0 : 2.0;
    v = scale(h * inv, y[static_cast<long>(c) * N + k]);
  }
  out[idx] = v;
}

// Per block of samples, the largest |a_c(t)|^2 of each column and of sum_c |a_c(t)|^2;
// blocks hold kStatThreads samples. A maximum does not depend on its order.
__global__ void __launch_bounds__(kStatThreads)
    stat_peak_blocks(const cplx* a, int C, int n_t, double* peaks) {
  __shared__ double red[kStatThreads];
  const int t = blockIdx.x * kStatThreads + threadIdx.x;
  double total = 0.0;
  for (int c = 0; c <= C; ++c) {
    double v = 0.0;
    if (c < C) {
      if (t < n_t) {
        const cplx z = a[static_cast<long>(c) * n_t + t];
        v = z.x * z.x + z.y * z.y;
        total = total + v;
      }
    } else {
      v = total;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int width = kStatThreads / 2; width > 0; width >>= 1) {
      if (threadIdx.x < width) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + width]);
      __syncthreads();
    }
    if (threadIdx.x == 0) peaks[static_cast<long>(blockIdx.x) * (C + 1) + c] = red[0];
    __syncthreads();
  }
}

// Per block of samples, the largest on-axis power sum_rows |sum_c M[row, c] a_c(t)|^2,
// the columns summed in order (native/src/general/statistics.cpp:517-603).
__global__ void __launch_bounds__(kStatThreads)
    stat_intensity_blocks(const cplx* a, int C, int n_t, const double* M, int rows,
                          double* peaks) {
  __shared__ double red[kStatThreads];
  const int t = blockIdx.x * kStatThreads + threadIdx.x;
  double power = 0.0;
  if (t < n_t) {
    for (int row = 0; row < rows; ++row) {
      double re = 0.0, im = 0.0;
      for (int c = 0; c < C; ++c) {
        const double m = M[row * C + c];
        const cplx z = a[static_cast<long>(c) * n_t + t];
        re = re + m * z.x;
        im = im + m * z.y;
      }
      power = power + (re * re + im * im);
    }
  }
  red[threadIdx.x] = power;
  __syncthreads();
  for (int width = kStatThreads / 2; width > 0; width >>= 1) {
    if (threadIdx.x < width) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + width]);
    __syncthreads();
  }
  if (threadIdx.x == 0) peaks[blockIdx.x] = red[0];
}

// The total power trace: sum_c |a_c(t)|^2 in column order at each sample.
__global__ void stat_total_power(const cplx* a, int C, int n_t, double* trace) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_t) return;
  double sum = 0.0;
  for (int c = 0; c < C; ++c) {
    const cplx z = a[static_cast<long>(c) * n_t + t];
    sum = sum + (z.x * z.x + z.y * z.y);
  }
  trace[t] = sum;
}

// The monitor's five sums, one partial set a block: each thread takes samples t in
// order of its stride, P(t) summed over the columns in order, then a fixed tree.
__global__ void __launch_bounds__(kMonitorThreads)
    monitor_partials(const double* samples, int layout, int C, int n_t, const double* columns,
                     const double* weights, double h, const double* h_at, int mode,
                     double* partials) {
  __shared__ double red[kMonitorThreads];
  const double step = h_at ? *h_at : h;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = blockIdx.x * kMonitorThreads + threadIdx.x; t < n_t;
       t += gridDim.x * kMonitorThreads) {
    double p = 0.0;
    for (int c = 0; c < C; ++c) {
      double v;
      if (layout == kSamplesReal) {
        const double x = samples[static_cast<long>(c) * n_t + t];
        v = x * x;
      } else {
        const double x = samples[2 * (static_cast<long>(c) * n_t + t)];
        const double y = samples[2 * (static_cast<long>(c) * n_t + t) + 1];
        v = layout == kSamplesComplex ? x * x + y * y : x * x;
      }
      p = p + (columns ? columns[c] * v : v);
    }
    const double rate = weights[3L * n_t + t];
    if (mode == kMonitorReading) {
      acc[0] += p;
      acc[1] += weights[t] * p;
      acc[2] += weights[static_cast<long>(n_t) + t] * p;
      acc[3] += weights[2L * n_t + t] * p;
      acc[4] += rate * p;
    } else {
      const double half = p * exp(-0.5 * rate * step);
      acc[0] += p * -expm1(-rate * step);
      acc[1] += half;
      acc[2] += weights[t] * half;
      acc[3] += weights[static_cast<long>(n_t) + t] * half;
      acc[4] += weights[2L * n_t + t] * half;
    }
  }
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    red[threadIdx.x] = acc[j];
    __syncthreads();
    for (int width = kMonitorThreads / 2; width > 0; width >>= 1) {
      if (threadIdx.x < width) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + width];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[5L * blockIdx.x + j] = red[0];
    __syncthreads();
  }
}

// The blocks' partials summed in block order by one thread: the reading written, or
// the split's removal added to the running total and the found energies stored.
__global__ void monitor_finish(const double* partials, int blocks, int mode, double* acc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double total[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < blocks; ++b)
#pragma unroll
    for (int j = 0; j < 5; ++j) total[j] = total[j] + partials[5L * b + j];
  if (mode == kMonitorReading) {
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] = total[j];
  } else {
    acc[0] = acc[0] + total[0];
#pragma unroll
    for (int j = 1; j < 5; ++j) acc[j] = total[j];
  }
}

// One workgroup: the densest node by a serial scan in node order (the first maximum),
// then the largest rate of its row by a fixed tree.
__global__ void __launch_bounds__(kStatThreads)
    plasma_densest(const double* frac, const double* rate, int Q, int T, double* out) {
  __shared__ double red[kStatThreads];
  __shared__ int densest;
  __shared__ double largest;
  if (threadIdx.x == 0) {
    int best = 0;
    double value = frac[static_cast<long>(T) - 1];
    for (int q = 1; q < Q; ++q) {
      const double f = frac[static_cast<long>(q) * T + T - 1];
      if (f > value) {
        value = f;
        best = q;
      }
    }
    densest = best;
    largest = value;
  }
  __syncthreads();
  double peak = 0.0;
  for (int t = threadIdx.x; t < T; t += kStatThreads)
    peak = fmax(peak, rate[static_cast<long>(densest) * T + t]);
  red[threadIdx.x] = peak;
  __syncthreads();
  for (int width = kStatThreads / 2; width > 0; width >>= 1) {
    if (threadIdx.x < width) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + width]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = largest;
    out[1] = red[0];
  }
}

// One workgroup a node: its end-of-window fraction and its largest rate by a fixed tree.
__global__ void __launch_bounds__(kStatThreads)
    plasma_populations(const double* frac, const double* rate, int T, double* out) {
  __shared__ double red[kStatThreads];
  const long q = blockIdx.x;
  double peak = 0.0;
  for (int t = threadIdx.x; t < T; t += kStatThreads) peak = fmax(peak, rate[q * T + t]);
  red[threadIdx.x] = peak;
  __syncthreads();
  for (int width = kStatThreads / 2; width > 0; width >>= 1) {
    if (threadIdx.x < width) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + width]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * q] = frac[q * T + T - 1];
    out[2 * q + 1] = red[0];
  }
}

}  // namespace

void launch_plasma_populations(const double* frac, const double* rate, int Q, int T,
                               double* out, hipStream_t s) {
  SP_CHECK(frac && rate && out && Q >= 1 && T >= 1, "the plasma column's arrays");
  hipLaunchKernelGGL(plasma_populations, dim3(Q), dim3(kStatThreads), 0, s, frac, rate, T, out);
  SP_HIP(hipGetLastError());
}

void launch_plasma_densest(const double* frac, const double* rate, int Q, int T, double* out,
                           hipStream_t s) {
  SP_CHECK(frac && rate && out && Q >= 1 && T >= 1, "the plasma column's arrays");
  hipLaunchKernelGGL(plasma_densest, dim3(1), dim3(kStatThreads), 0, s, frac, rate, Q, T, out);
  SP_HIP(hipGetLastError());
}

void launch_monitor_sums(const double* samples, MonitorSamples layout, int C, int n_t,
                         const double* columns, const double* weights, double h,
                         const double* h_at, MonitorMode mode, double* partials, double* acc,
                         hipStream_t s) {
  SP_CHECK(samples && weights && partials && acc, "the monitor's buffers");
  const int blocks = monitor_blocks(n_t);
  hipLaunchKernelGGL(monitor_partials, dim3(blocks), dim3(kMonitorThreads), 0, s, samples,
                     static_cast<int>(layout), C, n_t, columns, weights, h, h_at,
                     static_cast<int>(mode), partials);
  SP_HIP(hipGetLastError());
  hipLaunchKernelGGL(monitor_finish, dim3(1), dim3(64), 0, s, partials, blocks,
                     static_cast<int>(mode), acc);
  SP_HIP(hipGetLastError());
}

MonitorBuffers::~MonitorBuffers() {
  if (weights_) (void)hipFree(weights_);
  if (columns_) (void)hipFree(columns_);
  if (partials_) (void)hipFree(partials_);
  if (acc_) (void)hipFree(acc_);
}

void MonitorBuffers::bind(const double* weights, const double* columns, int C, int n_t,
                          hipStream_t s) {
  SP_CHECK(weights, "the monitor's weights are missing");
  if (!weights_ || C != C_ || n_t != n_t_) {
    if (weights_) (void)hipFree(weights_);
    if (columns_) (void)hipFree(columns_);
    if (partials_) (void)hipFree(partials_);
    if (acc_) (void)hipFree(acc_);
    C_ = C;
    n_t_ = n_t;
    SP_HIP(hipMalloc(&weights_, 4UL * n_t * sizeof(double)));
    SP_HIP(hipMalloc(&columns_, static_cast<std::size_t>(C) * sizeof(double)));
    SP_HIP(hipMalloc(&partials_, 5UL * monitor_blocks(n_t) * sizeof(double)));
    SP_HIP(hipMalloc(&acc_, 10 * sizeof(double)));
  }
  SP_HIP(hipMemcpyAsync(weights_, weights, 4UL * n_t * sizeof(double), hipMemcpyHostToDevice, s));
  has_columns_ = columns != nullptr;
  if (columns)
    SP_HIP(hipMemcpyAsync(columns_, columns, static_cast<std::size_t>(C) * sizeof(double),
                          hipMemcpyHostToDevice, s));
  SP_HIP(hipMemsetAsync(acc_, 0, 10 * sizeof(double), s));
  SP_HIP(hipStreamSynchronize(s));
  bound_ = true;
}

void MonitorBuffers::drain(double* out, hipStream_t s) {
  SP_HIP(hipMemcpyAsync(out, acc_, 10 * sizeof(double), hipMemcpyDeviceToHost, s));
  SP_HIP(hipMemsetAsync(acc_ + 5, 0, sizeof(double), s));
  SP_HIP(hipStreamSynchronize(s));
}

void PlaneStatistics::reading(const cplx* y, const double* columns, const double* weights,
                              double* partials, double* acc, hipStream_t s) {
  // The coarse samples: the inverse transform (an envelope) or the real part of the
  // analytic signal (a field, whose real part is the field itself).
  const std::size_t ct = static_cast<std::size_t>(C_) * n_t_;
  hipLaunchKernelGGL(stat_analytic_pack, dim3((ct + kStatThreads - 1) / kStatThreads),
                     dim3(kStatThreads), 0, s, y, C_, N_, n_t_, field_ ? 1 : 0, omega_,
                     nullptr, time_);
  SP_HIP(hipGetLastError());
  plan_.exec(time_, scratch_, +1, s);
  launch_monitor_sums(reinterpret_cast<const double*>(time_),
                      field_ ? kSamplesRealPart : kSamplesComplex, C_, n_t_, columns, weights,
                      0.0, nullptr, kMonitorReading, partials, acc, s);
}

PlaneStatistics::PlaneStatistics(int C, int N, int n_t, bool field, double dt,
                                 const double* omega)
    : C_(C), N_(N), n_t_(n_t), field_(field), dt_(dt), omega_(omega) {
  SP_CHECK(C >= 1 && N >= 2 && n_t >= 16 && (n_t & (n_t - 1)) == 0,
           "plane statistics need a power-of-two coarse grid of at least 16 samples");
  SP_CHECK(field ? N <= n_t / 2 + 1 : N == n_t, "the stored bins do not fit the coarse grid");
  const std::size_t ct = static_cast<std::size_t>(C) * n_t;
  SP_HIP(hipMalloc(&time_, ct * sizeof(cplx)));
  SP_HIP(hipMalloc(&scratch_, ct * sizeof(cplx)));
  SP_HIP(hipMalloc(&windows_, 2 * kMaxStatWindows * sizeof(double)));
  SP_HIP(hipMalloc(&sums_, static_cast<std::size_t>(C) * kSums * sizeof(double)));
  peak_blocks_ = (n_t + kStatThreads - 1) / kStatThreads;
  SP_HIP(hipMalloc(&peaks_, static_cast<std::size_t>(peak_blocks_) * (C + 1) * sizeof(double)));
  SP_HIP(hipMalloc(&trace_, static_cast<std::size_t>(n_t) * sizeof(double)));
  SP_HIP(hipMalloc(&on_axis_, 2 * static_cast<std::size_t>(C) * sizeof(double)));
  SP_HIP(hipMalloc(&intensity_, static_cast<std::size_t>(peak_blocks_) * sizeof(double)));
  SP_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_),
                       sizeof(double) * (static_cast<std::size_t>(C) * kSums +
                                         static_cast<std::size_t>(peak_blocks_) * (C + 1))));
  bytes_ = 2 * ct * sizeof(cplx) + static_cast<std::size_t>(C) * kSums * sizeof(double) +
           static_cast<std::size_t>(n_t) * sizeof(double) +
           static_cast<std::size_t>(peak_blocks_) * (C + 1) * sizeof(double);
  plan_.init(n_t, C);
}

PlaneStatistics::~PlaneStatistics() {
  if (time_) (void)hipFree(time_);
  if (scratch_) (void)hipFree(scratch_);
  if (windows_) (void)hipFree(windows_);
  if (sums_) (void)hipFree(sums_);
  if (peaks_) (void)hipFree(peaks_);
  if (trace_) (void)hipFree(trace_);
  if (on_axis_) (void)hipFree(on_axis_);
  if (intensity_) (void)hipFree(intensity_);
  if (host_) (void)hipHostFree(host_);
}

std::vector<double> PlaneStatistics::run(const cplx* y, std::uint64_t flags,
                                         const double* windows, int n_windows, hipStream_t s,
                                         std::vector<double>* trace, const StatExtras* extras,
                                         double* intensity) {
  SP_CHECK((flags & ~kStatBuilt) == 0,
           "statistics built: omega0, energy, energy_window, peak_power, peak_power_window, "
           "peak_power_total, fwhm, peak_intensity");
  const bool wants_intensity = flags & kStatPeakIntensity;
  if (wants_intensity) {
    SP_CHECK(extras && intensity && extras->on_axis && (extras->rows == 1 || extras->rows == 2),
             "on-axis synthesis must have one or two rows and n_modes columns");
    SP_CHECK(std::isfinite(extras->intensity_factor), "the intensity factor must be finite");
  }
  SP_CHECK(n_windows >= 0 && n_windows <= kMaxStatWindows, "at most 16 statistics windows");
  SP_CHECK(n_windows == 0 || windows, "the statistics windows are missing");
  for (int j = 0; j < n_windows; ++j)
    SP_CHECK(!(windows[2 * j] > windows[2 * j + 1]), "a statistics window's lo exceeds its hi");
  const bool spectral = flags & (kStatOmega0 | kStatEnergy | kStatEnergyWindow);
  const bool peaks = flags & (kStatPeakPower | kStatPeakPowerTotal | kStatFwhm | kStatPeakIntensity);
  SP_CHECK(!(flags & kStatFwhm) || trace, "the power trace needs somewhere to go");
  const int used_windows = (flags & kStatEnergyWindow) ? n_windows : 0;
  const std::size_t n_sums = static_cast<std::size_t>(C_) * kSums;
  if (spectral) {
    if (used_windows)
      SP_HIP(hipMemcpyAsync(windows_, windows, 2 * used_windows * sizeof(double),
                            hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(stat_spectral_sums, dim3(C_), dim3(kStatThreads), 0, s, y, N_, n_t_,
                       field_ ? 1 : 0, dt_, omega_, windows_, used_windows, sums_);
    SP_HIP(hipGetLastError());
    SP_HIP(hipMemcpyAsync(host_, sums_, n_sums * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (peaks) {
    const std::size_t ct = static_cast<std::size_t>(C_) * n_t_;
    hipLaunchKernelGGL(stat_analytic_pack, dim3((ct + kStatThreads - 1) / kStatThreads),
                       dim3(kStatThreads), 0, s, y, C_, N_, n_t_, field_ ? 1 : 0, omega_,
                     nullptr, time_);
    SP_HIP(hipGetLastError());
    plan_.exec(time_, scratch_, +1, s);
    hipLaunchKernelGGL(stat_peak_blocks, dim3(peak_blocks_), dim3(kStatThreads), 0, s, time_,
                       C_, n_t_, peaks_);
    SP_HIP(hipGetLastError());
    SP_HIP(hipMemcpyAsync(host_ + n_sums, peaks_,
                          static_cast<std::size_t>(peak_blocks_) * (C_ + 1) * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    if (wants_intensity) {
      SP_HIP(hipMemcpyAsync(on_axis_, extras->on_axis,
                            static_cast<std::size_t>(extras->rows) * C_ * sizeof(double),
                            hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(stat_intensity_blocks, dim3(peak_blocks_), dim3(kStatThreads), 0, s,
                         time_, C_, n_t_, on_axis_, extras->rows, intensity_);
      SP_HIP(hipGetLastError());
    }
    if (flags & kStatFwhm) {
      hipLaunchKernelGGL(stat_total_power, dim3((n_t_ + kStatThreads - 1) / kStatThreads),
                         dim3(kStatThreads), 0, s, time_, C_, n_t_, trace_);
      SP_HIP(hipGetLastError());
      trace->assign(n_t_, 0.0);
      SP_HIP(hipMemcpyAsync(trace->data(), trace_, static_cast<std::size_t>(n_t_) * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    }
  }
  // Each window's gated spectrum through its own transform: the peak of every column.
  std::vector<double> window_peaks;
  if ((flags & kStatPeakPowerWindow) && n_windows > 0) {
    if (!(flags & kStatEnergyWindow))
      SP_HIP(hipMemcpyAsync(windows_, windows, 2 * n_windows * sizeof(double),
                            hipMemcpyHostToDevice, s));
    const std::size_t ct = static_cast<std::size_t>(C_) * n_t_;
    const std::size_t per = static_cast<std::size_t>(peak_blocks_) * (C_ + 1);
    std::vector<double> blocks(per);
    for (int j = 0; j < n_windows; ++j) {
      hipLaunchKernelGGL(stat_analytic_pack, dim3((ct + kStatThreads - 1) / kStatThreads),
                         dim3(kStatThreads), 0, s, y, C_, N_, n_t_, field_ ? 1 : 0, omega_,
                         windows_ + 2 * j, time_);
      SP_HIP(hipGetLastError());
      plan_.exec(time_, scratch_, +1, s);
      hipLaunchKernelGGL(stat_peak_blocks, dim3(peak_blocks_), dim3(kStatThreads), 0, s, time_,
                         C_, n_t_, peaks_);
      SP_HIP(hipGetLastError());
      SP_HIP(hipMemcpyAsync(blocks.data(), peaks_, per * sizeof(double), hipMemcpyDeviceToHost,
                            s));
      SP_HIP(hipStreamSynchronize(s));
      for (int c = 0; c < C_; ++c) {
        double m = 0.0;
        for (int b = 0; b < peak_blocks_; ++b) m = std::fmax(m, blocks[b * (C_ + 1) + c]);
        window_peaks.push_back(m);
      }
    }
  }
  std::vector<double> intensity_blocks;
  if (wants_intensity) {
    intensity_blocks.assign(peak_blocks_, 0.0);
    SP_HIP(hipMemcpyAsync(intensity_blocks.data(), intensity_,
                          intensity_blocks.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  SP_HIP(hipStreamSynchronize(s));
  if (wants_intensity) {
    double peak = 0.0;
    for (double v : intensity_blocks) peak = std::fmax(peak, v);
    *intensity = extras->intensity_factor * peak;
  }
  std::vector<double> out;
  const double* sums = host_;
  if (flags & kStatOmega0) {
    // An empty column (its density at most one eps of the largest column's) has no
    // centroid (definitions.py:340-355).
    double largest = 0.0;
    for (int c = 0; c < C_; ++c) largest = std::fmax(largest, sums[c * kSums + 1]);
    const double floor = std::numeric_limits<double>::epsilon() * largest;
    for (int c = 0; c < C_; ++c) {
      const double total = sums[c * kSums + 1];
      out.push_back(total > floor ? sums[c * kSums + 2] / total
                                  : std::numeric_limits<double>::quiet_NaN());
    }
  }
  if (flags & kStatEnergy)
    for (int c = 0; c < C_; ++c) out.push_back(sums[c * kSums]);
  if (flags & kStatEnergyWindow)
    for (int j = 0; j < n_windows; ++j)
      for (int c = 0; c < C_; ++c) out.push_back(sums[c * kSums + 3 + j]);
  const double* blocks = host_ + n_sums;
  auto peak_of = [&](int c) {
    double m = 0.0;
    for (int b = 0; b < peak_blocks_; ++b) m = std::fmax(m, blocks[b * (C_ + 1) + c]);
    return m;
  };
  if (flags & kStatPeakPower)
    for (int c = 0; c < C_; ++c) out.push_back(peak_of(c));
  out.insert(out.end(), window_peaks.begin(), window_peaks.end());
  if (flags & kStatPeakPowerTotal) out.push_back(peak_of(C_));
  return out;
}

}  // namespace sp
