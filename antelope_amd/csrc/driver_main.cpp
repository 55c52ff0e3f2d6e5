// antelope_mi355x - the reference's executable (antelope.cpp:50-593) on one
// MI355X: reads inputParam.cfg from the working directory (or the path
// given), propagates the whole k grid on the GPU and writes the reference's
// output files into the working directory (or --outdir), snapshots included
// when calc.shotNumber > 0. A config whose name ends in .atl is read in the
// .atl format of src/inputFile.atl. --hhg FILE also writes the HHG spectrum
// of the total current (pulse 0's w0, x and y axes) in the layout of the
// reference's post-processing script (AnalysisReader_PyDir.py:367-382).
//
//   antelope_mi355x [config] [--outdir DIR] [--hhg FILE] [--quiet]
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "antelope/sbe_c.h"
#include "antelope/sbe_frontend.h"

int main(int argc, char **argv)
{
    std::string path = "inputParam.cfg", outdir = ".", hhg;  // antelope.cpp:93
    bool quiet = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--outdir") && i + 1 < argc) outdir = argv[++i];
        else if (!std::strcmp(argv[i], "--hhg") && i + 1 < argc) hhg = argv[++i];
        else if (!std::strcmp(argv[i], "--quiet")) quiet = true;
        else path = argv[i];
    }
    sbe_spec *spec = nullptr;
    const bool atl = path.size() > 4 && path.compare(path.size() - 4, 4, ".atl") == 0;
    if ((atl ? sbe_spec_parse_atl_file(path.c_str(), &spec) : sbe_spec_parse_file(path.c_str(), &spec)) != SBE_OK) {
        std::fprintf(stderr, "%s\n", atl ? sbe_atl_last_error() : sbe_frontend_last_error());
        return 1;
    }
    if (*sbe_spec_warnings(spec)) std::fprintf(stderr, "Warning: %s", sbe_spec_warnings(spec));
    const sbe_config *cfg = nullptr;
    const sbe_model *model = nullptr;
    sbe_spec_get(spec, &cfg, &model);
    sbe_ctx *ctx = nullptr;
    if (sbe_create(cfg, model, -1, -1, &ctx) != SBE_OK) {
        std::fprintf(stderr, "%s\n", sbe_last_error(nullptr));
        sbe_spec_free(spec);
        return 1;
    }
    sbe_info info;
    sbe_get_info(ctx, &info);
    if (!quiet)
        std::printf("Nk = %lld, Nt = %lld, dt = %g, atmin = %g, dV = %.6e\n", (long long)info.nk_total,
                    (long long)info.nt, info.dt, info.atmin, info.dV);
    const int64_t chunk = info.nt / 10 > 0 ? info.nt / 10 : 1;
    // snapshots every `shot` steps, taken before the step (antelope.cpp:340-421)
    const int64_t shot = sbe_snapshot_interval(cfg, sbe_spec_shot_number(spec));
    std::vector<double> band;
    if (shot > 0) band.resize(static_cast<size_t>(info.nk_total) * info.nband * info.nband * 2);
    for (int64_t s = 0; s < info.nt;) {
        if (shot > 0 && s % shot == 0) {
            if (sbe_get_state(ctx, SBE_BASIS_BAND, band.data()) != SBE_OK ||
                sbe_write_snapshot(outdir.c_str(), cfg, model, band.data(), s > 0) != SBE_OK) {
                std::fprintf(stderr, "snapshot: %s%s\n", sbe_last_error(ctx), sbe_frontend_last_error());
                sbe_destroy(ctx);
                sbe_spec_free(spec);
                return 1;
            }
        }
        int64_t e = s + chunk < info.nt ? s + chunk : info.nt;
        if (shot > 0 && (s / shot + 1) * shot < e) e = (s / shot + 1) * shot;
        if (sbe_run(ctx, s, e) != SBE_OK || sbe_synchronize(ctx) != SBE_OK) {
            std::fprintf(stderr, "%s\n", sbe_last_error(ctx));
            sbe_destroy(ctx);
            sbe_spec_free(spec);
            return 1;
        }
        if (!quiet && (e / chunk != s / chunk || e == info.nt)) std::printf("%lld / %lld\n", (long long)e, (long long)info.nt);
        s = e;
    }
    const int N = info.nband;
    std::vector<double> ji(3 * info.nt), ja(3 * info.nt), rb(2 * N * N * info.nt), bp(N * info.nt);
    int rc = sbe_get_series(ctx, ji.data(), ja.data(), rb.data(), bp.data());
    if (rc != SBE_OK) std::fprintf(stderr, "%s\n", sbe_last_error(ctx));
    sbe_destroy(ctx);
    if (rc == SBE_OK) {
        std::printf("\n\nWriting output-data\n\n");
        rc = sbe_write_outputs(outdir.c_str(), cfg, model, &info, ji.data(), ja.data(), rb.data(), bp.data());
        if (rc != SBE_OK) std::fprintf(stderr, "%s\n", sbe_frontend_last_error());
    }
    if (rc == SBE_OK && !hhg.empty()) {
        // the script reads the .dat files, which hold the sums x dV
        for (double &v : ji) v *= info.dV;
        for (double &v : ja) v *= info.dV;
        rc = cfg->npulses > 0 ? sbe_write_hhg(hhg.c_str(), info.nt, ji.data(), ja.data(), info.dt, cfg->pulses[0].w0, 2)
                              : SBE_ERR_CONFIG;
        if (rc != SBE_OK) std::fprintf(stderr, "%s\n", sbe_spectrum_last_error());
    }
    sbe_spec_free(spec);
    return rc == SBE_OK ? 0 : 1;
}
