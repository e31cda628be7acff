// TimeSeriesProcessor::run of the reference (time_series_processor.cpp:
// 21-210): for every snapshot of the series, the eigen-decomposition of the
// dense Hessian and the speeds on the device, then one gzip-compressed text
// record in the reference's column order and number formatting (a default
// std::ostream: six significant digits):
//   time, eigenvalues, participation numbers, s_i^2 / lambda_i, speeds, x,
//   s_i = sum_j v_ij b_j, the first numberOfEigenVecToWrite eigenvectors and,
//   with calculateDerivativeEVAnal, the eigenvalue derivatives (zeros unless
//   the protocol is fixed-rate-stress, time_series_processor.cpp:92-111)
// or, with writeCorrelMatrices = W > 0, the first three groups followed per
// mode by the three W x W histograms of v_i^2 v_j^2 over the pair's (dx, dy)
// (same-sign pairs, opposite-sign pairs, all pairs).
#include "../../include/sdddst/time_series_processor.h"

#include <zlib.h>

#include <cmath>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sdddst/utility.h"
#include "../../include/sdddst_amd.h"

namespace sdddstEV {

namespace {

void check(int status, const char *what)
{
    if (status != SDDDST_OK) throw std::runtime_error(std::string(what) + ": " + sdddst_last_error());
}

// One device context for one snapshot.
struct Ctx {
    sdddst_ctx *h = nullptr;
    ~Ctx()
    {
        if (h) sdddst_ctx_destroy(h);
    }
};

void write(gzFile f, const std::string &s)
{
    if (!s.empty() && gzwrite(f, s.data(), (unsigned)s.size()) != (int)s.size())
        throw std::runtime_error("EV analysis: writing the result file failed");
}

}  // namespace

TimeSeriesProcessor::TimeSeriesProcessor(std::shared_ptr<sdddstCore::SimulationData> simData,
                                         std::shared_ptr<DataTimeSeries> dateTimeSeries,
                                         sdddstCore::StressProtocol *stressProtocol)
    : sD(simData), dTS(dateTimeSeries)
{
    sD->externalStressProtocol.reset(stressProtocol);
}

void TimeSeriesProcessor::run()
{
    if (!dTS) throw std::runtime_error("EV analysis: --file-list and --pd-configuration are required");
    gzFile f = gzopen(sD->eVAnalResultFile.c_str(), "wb");
    if (!f) throw std::runtime_error("EV analysis: cannot open " + sD->eVAnalResultFile);
    try {
        while (dTS->next(sD)) {
            const int n = (int)sD->dc, pc = (int)sD->pc;
            sD->externalStressProtocol->calculateStress(sD->simTime, sD->dislocations,
                                                        sdddstCore::StressProtocolStepType::Original);
            std::cout << sD->simTime << "\n";
            std::vector<double> x(n), y(n), b(n), px(pc), py(pc);
            for (int i = 0; i < n; ++i) {
                x[i] = sD->dislocations[i].x;
                y[i] = sD->dislocations[i].y;
                b[i] = sD->dislocations[i].b;
            }
            for (int i = 0; i < pc; ++i) {
                px[i] = sD->points[i].x;
                py[i] = sD->points[i].y;
            }
            Ctx ctx;
            sdddst_params p{sD->cutOffMultiplier, sD->prec, sD->A, sD->KASQR, (int)sD->ic, sD->fieldMode};
            check(sdddst_ctx_create(n, pc, &p, &ctx.h), "sdddst_ctx_create");
            check(sdddst_upload_state(ctx.h, 0, x.data(), y.data(), b.data(), n), "upload");
            if (pc > 0) check(sdddst_upload_points(ctx.h, px.data(), py.data(), pc), "upload points");
            std::vector<double> lam(n), vec((size_t)n * n), pn(n), gan(n), sum(n), speeds, deriv;
            const bool plain = sD->writeCorrelMatrices == 0;
            if (plain) {
                speeds.resize(n);
                check(sdddst_speeds(ctx.h, 0, SDDDST_V_SCRATCH,
                                    sD->externalStressProtocol->getStress(sdddstCore::StressProtocolStepType::Original),
                                    0, speeds.data()),
                      "speeds");
            }
            // the derivatives need the fixed-rate protocol's sigma'; other
            // protocols write zeros
            const bool with_deriv = plain && sD->calculateDerivativeEVAnal &&
                                    sD->externalStressProtocol->getType() == "fixed-rate-stress";
            if (plain && sD->calculateDerivativeEVAnal) deriv.assign(n, 0.0);
            check(sdddst_ev_analysis(ctx.h, 0, with_deriv ? speeds.data() : nullptr,
                                     with_deriv ? sD->externalStressProtocol->getStressDerivative(sD->simTime) : 1.0,
                                     lam.data(), vec.data(), pn.data(), gan.data(), sum.data(),
                                     with_deriv ? deriv.data() : nullptr),
                  "EV decomposition");
            auto v = [&](int mode, int j) { return vec[(size_t)mode * n + j]; };
            std::ostringstream out;
            out << sD->simTime;
            for (int i = 0; i < n; ++i) out << " " << lam[i];   // EV
            for (int i = 0; i < n; ++i) out << " " << pn[i];    // PN
            for (int i = 0; i < n; ++i) out << " " << gan[i];   // 1/Gael (GAN)
            if (plain) {
                for (int i = 0; i < n; ++i) out << " " << speeds[i];   // speeds
                for (int i = 0; i < n; ++i) out << " " << x[i];        // X coordinates
                for (int i = 0; i < n; ++i) out << " " << sum[i];      // SUM
                for (int i = 0; i < sD->numberOfEigenVecToWrite && i < n; ++i)
                    for (int j = 0; j < n; ++j) out << " " << v(i, j);
                for (const double d : deriv) out << " " << d;         // eigenvalue derivatives
                out << "\n";
                write(f, out.str());
                continue;
            }
            out << "\n";
            write(f, out.str());
            const int W = sD->writeCorrelMatrices;
            std::vector<double> pp((size_t)W * W), pnm((size_t)W * W), all((size_t)W * W);
            for (int mode = 0; mode < n; ++mode) {
                std::fill(pp.begin(), pp.end(), 0.0);
                std::fill(pnm.begin(), pnm.end(), 0.0);
                std::fill(all.begin(), all.end(), 0.0);
                for (int i = 0; i < n; ++i) {
                    for (int j = 0; j < n; ++j) {
                        if (i == j) continue;
                        double dx = x[i] - x[j], dy = y[i] - y[j];
                        normalize(dx);
                        normalize(dy);
                        const size_t cell = (size_t)(int)floor((dx + 0.5) * W) * W + (size_t)(int)floor((dy + 0.5) * W);
                        const double w = pow(v(mode, j), 2.0) * pow(v(mode, i), 2.0);
                        if (b[i] != b[j]) pnm[cell] += w;
                        if (b[i] == b[j]) pp[cell] += w;
                        all[cell] += w;
                    }
                }
                std::ostringstream m;
                for (const std::vector<double> *a : {&pp, &pnm, &all})
                    for (int r = 0; r < W; ++r) {
                        for (int c = 0; c < W; ++c) m << (*a)[(size_t)r * W + c] << " ";
                        m << "\n";
                    }
                write(f, m.str());
            }
        }
    } catch (...) {
        gzclose(f);
        throw;
    }
    if (gzclose(f) != Z_OK) throw std::runtime_error("EV analysis: closing the result file failed");
}

}  // namespace sdddstEV
