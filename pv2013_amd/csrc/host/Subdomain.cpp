#include "Subdomain.h"

#include <cstring>

SubDomain::SubDomain(Geometry& g, const CaseConfig& cfg, lbm_ctx* ctx) : Geometry_(g) {
    my_Nx_ = g.GeomSize[0];
    my_Ny_ = g.GeomSize[1];
    my_Nz_ = g.GeomSize[2];
    lbm_domain_desc d;
    std::memset(&d, 0, sizeof d);
    d.dims[0] = my_Nx_;
    d.dims[1] = my_Ny_;
    d.dims[2] = my_Nz_;
    d.q = cfg.q;
    d.storage = cfg.storage;
    d.scheme = cfg.scheme;
    d.quirks = cfg.quirks;
    d.tau = cfg.tau;
    d.flags = g.flags.data();
    d.bcs = cfg.bcs.data();
    d.nbc = (int32_t)cfg.bcs.size();
    d.fast_axis = -1;
    if (cfg.autotune) {
        lbm_tune_result t;
        if (lbm_autotune(ctx, &d, 128, cfg.autotune_cache.empty() ? nullptr : cfg.autotune_cache.c_str(), &t) !=
            LBM_OK) {
            ok = false;
            error = lbm_last_error(ctx);
            return;
        }
        d.scheme = t.scheme;
        d.fast_axis = t.fast_axis;
        d.options = t.options;
        d.block = t.block;
        tuned = t;
    }
    if (lbm_domain_create(ctx, &d, &dom) != LBM_OK) {
        ok = false;
        error = lbm_last_error(ctx);
        return;
    }
    int64_t bytes = 0;
    lbm_domain_info(dom, &fluid_nodes, &boundary_records, &bytes);
    for (int i = 0; i < my_Nx_; ++i)
        for (int j = 0; j < my_Ny_; ++j)
            for (int k = 0; k < my_Nz_; ++k) {
                const int64_t s = i + (int64_t)my_Nx_ * (j + (int64_t)my_Ny_ * k);
                if (g.flags[s] == 2) wall_sites.push_back(s);
            }
    for (const WaveSpec& w : cfg.waves)
        if (lbm_set_waveform(dom, w.slot, w.kind, w.period, (int)w.a.size(), w.a.data(), w.phi.data(), w.dt) != LBM_OK) {
            ok = false;
            error = lbm_last_error(ctx);
            return;
        }
}

SubDomain::~SubDomain() { lbm_domain_destroy(dom); }
