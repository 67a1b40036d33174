// psa_trk.h - open-loop channel state shared by the tracking plans
// (psa_corr.hip: standard correlators, psa_l6.hip: L6D / L6E FFT tracking).
#pragma once
#include "psa_internal.h"

namespace psa {

// Per-channel parameters on the device.
struct TrkChDev {
    double fs, fi, fc_sig, T, fd0, drift, phi0, coff0;
};

// Epoch state of one (channel, epoch) of the standard-correlator plan:
// mixing parameters and the code offset in samples (cs = coff fs, the
// origin of the tap geometry).
struct TrkEpoch {
    int ix, nseg;
    uint32_t p0, p1, s;
    int idle;           // 1: no correlation (the channel has lost lock)
    int neg;            // 1: the taps are negated (a wiped secondary-code chip)
    double cs;
};

// sdr_mix_carr phase setup (src/sdr_func.c:776-778, 824-836) for N samples
// of a ring of buffN from index ix (0 <= ix <= buffN) at phase phi:
// first-segment phase p0, step s, the split nseg and the phase p1 there
// (the reference build fuses phi + step n into fma(n, step, phi)).
__device__ __forceinline__ void trk_mix_params(int buffN, int ix, int N, double fs, double fc, double phi,
                                               uint32_t &p0, uint32_t &p1, int &nseg, uint32_t &s)
{
#pragma clang fp contract(off)
    const double scale = 4294967296.0;
    const double step = fc / fs;
    p0 = (uint32_t)((phi - floor(phi)) * scale);
    s = (uint32_t)(int)(step * scale);
    if (ix + N <= buffN) {
        nseg = N;
        p1 = 0;
    } else {
        const int n = buffN - ix;
        const double phi2 = fma((double)n, step, phi);
        nseg = n;
        p1 = (uint32_t)((phi2 - floor(phi2)) * scale);
    }
}

// Open-loop NCO / code advance, src/sdr_ch.c:694-703 with adj_coff
// (674-689), in FP64, run by one 64-thread workgroup per channel.  Per chunk
// of 64 epochs the lanes compute the epoch-local terms (tau, fc = fi + fd
// and -fd / fc_sig) in parallel, lane 0 runs the order-dependent recurrence
// (phase wrap, code-offset wrap) over the chunk, and then lane l calls
// emit(e, phi, coff, fc) for epoch e = chunk start + l with the state after
// that epoch's update.  The reference build fuses coff -= fd / fc * tau and
// phi += fc * tau into fnmadd / fmadd; the same fused forms are written
// here explicitly.
#define TRK_SCHUNK 64
template <class Emit>
__device__ __forceinline__ void trk_open_loop(const TrkChDev &ch, int n_epoch, const Emit &emit)
{
#pragma clang fp contract(off)
    __shared__ double s_a[TRK_SCHUNK], s_tau[TRK_SCHUNK], s_fc[TRK_SCHUNK], s_coff[TRK_SCHUNK], s_phi[TRK_SCHUNK];
    const int lane = threadIdx.x;
    double phi = ch.phi0, coff = ch.coff0;    // lane 0's running state
    for (int e0 = 0; e0 < n_epoch; e0 += TRK_SCHUNK) {
        const int e = e0 + lane;
        double fc = 0.0;
        if (e < n_epoch) {
            // t of this epoch and of the previous one ((double)e T, 0 at e = 0)
            const double t = (double)(e + 1) * ch.T;
            const double time = (double)e * ch.T;
            const double fd = ch.fd0 + ch.drift * t;
            fc = ch.fi + fd;
            s_a[lane] = -(fd / ch.fc_sig);
            s_tau[lane] = t - time;
            s_fc[lane] = fc;
        }
        __syncthreads();
        if (lane == 0) {
            const int ne = min(TRK_SCHUNK, n_epoch - e0);
            for (int i = 0; i < ne; i++) {
                const double tau = s_tau[i];
                coff = fma(s_a[i], tau, coff);
                phi = fma(s_fc[i], tau, phi);
                phi = phi - floor(phi);
                if (coff >= ch.T) coff = coff - ch.T;
                else if (coff < 0.0) coff = coff + ch.T;
                s_coff[i] = coff;
                s_phi[i] = phi;
            }
        }
        __syncthreads();
        if (e < n_epoch) emit(e, s_phi[lane], s_coff[lane], fc);
        __syncthreads();
    }
}

}  // namespace psa
