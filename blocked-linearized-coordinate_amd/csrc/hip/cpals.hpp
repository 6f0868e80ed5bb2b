// cpals.hpp -- the CP-ALS solver state of one tensor and the pieces of a
// sweep, shared by the single-tensor driver (cpals.hip) and the batched
// driver (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "blco_hip.h"
#include "device.hpp"

struct blco_cpals_state {
    blco_dev* d = nullptr;
    blco_cpals_opts o{};
    int N = 0, R = 0, ft = BLCO_F64, mt = BLCO_F64;
    hipStream_t st = nullptr;
    void* U[BLCO_MAX_MODES] = {};
    double* G[BLCO_MAX_MODES] = {};
    void* Mbuf = nullptr;
    double *lambda = nullptr, *inv = nullptr, *fac = nullptr, *work = nullptr;
    int* status = nullptr;
    double *gram_part = nullptr, *stats = nullptr, *ip = nullptr, *fitbuf = nullptr;
    // the factors are stored unnormalised: U[m] = (normalised factor) diag(lambda_m);
    // dinv[m * R + r] = 1 / lambda_m[r] (0 for a zero column; 1 before mode m's
    // first update), and colscale the product of the other modes' dinv for
    // the mode being updated
    double *dinv = nullptr, *colscale = nullptr;
    double* norm_part = nullptr;
    double* h_fit = nullptr;  // pinned
    double normX2 = 0.0;
    int order[BLCO_MAX_MODES] = {};
    int kern[BLCO_MAX_MODES] = {};
    hipEvent_t ev0[BLCO_MAX_MODES] = {}, ev1[BLCO_MAX_MODES] = {};
    // side stream of the sweep: mode n's R x R factorisation needs only the
    // other modes' Grams, so it runs there while mode n's MTTKRP runs on st
    hipStream_t side = nullptr;
    hipEvent_t fork[BLCO_MAX_MODES] = {}, join[BLCO_MAX_MODES] = {};
    hipEvent_t it0 = nullptr, it1 = nullptr;
    double mttkrp_ms[BLCO_MAX_MODES] = {};
    double iter_ms = 0.0;
    int iters_timed = 0;
    int iters_done = 0;
    double fit_old = 0.0;
    bool stopped = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace blco {

// Host factor conversions (f64 <-> the solver's factor type).
void to_dtype(const double* src, void* dst, uint64_t n, int dt);
void from_dtype(const void* src, double* dst, uint64_t n, int dt);
// The ALS step of mode n after its MTTKRP is in s->Mbuf: solve, lambda,
// normalisation and the new Gram (plus the fit's inner product when n is the
// last mode of the sweep), on stream st.
int enqueue_step(blco_cpals_state* s, int n, int which_norm, hipStream_t st);
// The two halves of enqueue_step: the factorisation of V_n (needs the other
// modes' Grams only) and the rest (needs it and the MTTKRP).
int enqueue_factorize(blco_cpals_state* s, int n, hipStream_t st);
int enqueue_update(blco_cpals_state* s, int n, int which_norm, hipStream_t st);
// The fit of the current model into s->fitbuf, on stream st.
int enqueue_fit(blco_cpals_state* s, hipStream_t st);

}  // namespace blco
