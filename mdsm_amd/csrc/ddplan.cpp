// De-dispersion plan generator: DM ranges (passes) with step and time
// downsampling chosen from the smearing budget, so plans are generated
// rather than hand-written.  Restates the algorithm of the reference's
// python/ddplan.py (observation, dedisp_method, dm_steps; :6-253).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mdsm_internal.h"

using namespace mdsm;

namespace {

const double kAllowDDM[] = {0.0001, 0.0002, 0.0005, 0.001, 0.002, 0.005, 0.01, 0.02, 0.03,
                            0.05,   0.1,    0.2,    0.3,   0.5,   1.0,   2.0,  3.0,  5.0,
                            10.0,   20.0,   30.0,   50.0,  100.0, 200.0, 300.0, 500.0, 1000.0};
constexpr int kNDDM = sizeof(kAllowDDM) / sizeof(kAllowDDM[0]);
constexpr int kNDown = 20;  // downsampling factors 2^0 .. 2^19

struct Obs {
    double dt, f_ctr, bw, chanwidth;
};

// smearing in ms of a DM over a bandwidth bw (MHz) centred at f (ddplan.py:126-132)
double dm_smear(double dm, double bw, double f) { return 1000.0 * dm * bw / (0.0001205 * std::pow(f, 3.0)); }
double bw_smear(double step, double bw, double f) { return dm_smear(0.5 * step, bw, f); }        // :134-141
double guess_step(double dt, double bw, double f) { return dt * 0.0001205 * std::pow(f, 3.0) / (0.5 * bw); }  // :143-149
double sub_smear(double substep, int nsub, double bw, double f)                                   // :151-161
{
    if (nsub == 0) return 0.0;
    return dm_smear(0.5 * substep, bw / nsub, f);
}

struct Method {  // dedisp_method (ddplan.py:23-63)
    double lodm, hidm, ddm, dsubdm;
    int downsamp, numdms, dms_per_call, ncalls;
};

Method make_method(const Obs &o, int downsamp, double lodm, double hidm, double ddm, int nsub)
{
    Method m{};
    m.lodm = lodm;
    m.ddm = ddm;
    m.downsamp = downsamp;
    const double bws = bw_smear(ddm, o.bw, o.f_ctr);
    double subs;
    int per = 0;
    if (nsub) {
        per = 2;
        for (;;) {
            const double next = sub_smear((per + 2) * ddm, nsub, o.bw, o.f_ctr);
            if (next > 0.8 * std::min(bws, 1000.0 * o.dt * downsamp)) {
                m.dsubdm = per * ddm;
                subs = sub_smear(m.dsubdm, nsub, o.bw, o.f_ctr);
                break;
            }
            per += 2;
        }
    } else {
        m.dsubdm = ddm;
        subs = sub_smear(m.dsubdm, nsub, o.bw, o.f_ctr);
    }
    // DM where one channel's smearing is twice all other smearing (:84-94)
    const double other = std::sqrt(std::pow(1000.0 * o.dt, 2.0) +
                                   std::pow(1000.0 * o.dt * downsamp, 2.0) + std::pow(bws, 2.0) +
                                   std::pow(subs, 2.0));
    double cross = 2.0 * 0.001 * other / o.chanwidth * 0.0001205 * std::pow(o.f_ctr, 3.0);
    if (cross > hidm) cross = hidm;
    m.numdms = (int)std::ceil((cross - lodm) / ddm);
    if (nsub) {
        m.ncalls = (int)std::ceil(m.numdms * ddm / m.dsubdm);
        m.numdms = m.ncalls * per;
        m.dms_per_call = per;
    } else {
        m.ncalls = 1;
        m.dms_per_call = m.numdms;
    }
    m.hidm = lodm + m.numdms * ddm;
    return m;
}

}  // namespace

extern "C" int mdsm_ddplan(const mdsm_ddplan_params *p, mdsm_pass *passes, uint32_t cap,
                           uint32_t *npasses)
{
    if (!p || !npasses || (cap && !passes)) return fail(MDSM_ERR_INVALID, "NULL argument");
    if (!(p->dt > 0.0) || !(p->f_ctr > 0.0) || !(p->bw > 0.0) || p->nchans < 1 ||
        !(p->hidm > p->lodm) || !(p->lodm >= 0.0))
        return fail(MDSM_ERR_INVALID, "ddplan: need dt, f_ctr, bw > 0, nchans >= 1, 0 <= lodm < hidm");
    const Obs o{p->dt, p->f_ctr, p->bw, p->bw / p->nchans};
    const int nsub = (int)p->nsubs;
    // dm_steps (ddplan.py:163-253)
    int ids = 0, idm = 0;
    int downsamp = 1;
    const double dtms = 1000.0 * o.dt;
    const double ff = 1.2;
    const double min_chan = dm_smear(p->lodm + 0.5 * kAllowDDM[0], o.chanwidth, o.f_ctr);
    const double min_bw = bw_smear(kAllowDDM[0], o.bw, o.f_ctr);
    const double ok = std::max({p->ok_smearing, min_chan, min_bw, dtms});
    if (ff * min_chan > dtms || ok > dtms) {
        const double okval = ok > ff * min_chan ? ok : ff * min_chan;
        while (ids + 1 < kNDown && dtms * std::ldexp(1.0, ids + 1) < okval) ++ids;
        downsamp = 1 << ids;
    }
    const double ddm0 = guess_step(o.dt * downsamp, o.bw, o.f_ctr);
    while (idm + 1 < kNDDM && kAllowDDM[idm + 1] < ff * ddm0) ++idm;
    std::vector<Method> ms;
    ms.push_back(make_method(o, downsamp, p->lodm, p->hidm, kAllowDDM[idm], nsub));
    while (ms.back().hidm < p->hidm) {
        if (ids + 1 >= kNDown) return fail(MDSM_ERR_INVALID, "ddplan: downsampling exhausted");
        ++ids;
        downsamp = 1 << ids;
        const double eff = dtms * downsamp;
        while (idm + 1 < kNDDM && bw_smear(kAllowDDM[idm + 1], o.bw, o.f_ctr) < ff * eff) ++idm;
        ms.push_back(make_method(o, downsamp, ms.back().hidm, p->hidm, kAllowDDM[idm], nsub));
        if (ms.size() > 64) return fail(MDSM_ERR_INVALID, "ddplan: more than 64 passes");
    }
    *npasses = (uint32_t)ms.size();
    if (cap < ms.size()) return fail(MDSM_ERR_CAPACITY, "ddplan: pass buffer too small");
    for (size_t i = 0; i < ms.size(); ++i) {
        const Method &m = ms[i];
        mdsm_pass &q = passes[i];
        q.lowdm = (float)m.lodm;
        q.highdm = (float)m.hidm;
        q.dmstep = (float)m.ddm;
        q.sub_dmstep = (float)m.dsubdm;
        q.binsize = (uint32_t)m.downsamp;
        q.ndms = (uint32_t)m.numdms;
        q.ncalls = (uint32_t)m.ncalls;
        q.calldms = (uint32_t)m.dms_per_call;
    }
    return MDSM_OK;
}
