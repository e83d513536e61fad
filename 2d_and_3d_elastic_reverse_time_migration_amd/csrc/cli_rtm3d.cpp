// cli_rtm3d.cpp - command-line driver for 3D elastic RTM on one GPU.
//
// Stands in for the reference's absent main (3D_elastic_RTM/makefile builds
// 3D_elastic_modeling.cu, which is not in the checkout).  It reads the same
// getpar keys (`rtm3d par=file key=value ...`), loads the raw-float32 models,
// and for every shot (shot_num_x * shot_num_y * shot_num_z) runs the
// reference's three loops (3D_forward_or_back_together.cu:9-453):
//   join_shot == 0: model the record in the true model,
//   join_shot != 0: read it from infile_shot_name ([3][lt][ry][rx] float32);
// then the source forward pass with excitation and the receiver backward
// pass with imaging, and writes ./result/result-{tp,pp,ps}-{sy}-{sx} and
// ./someoutput/cut-ex-*-{sy}-{sx} (3D_output_file.cu:155-250,309-428).
// A shot whose three result files already exist with their full size is
// skipped (resume); files are written under a temporary name and renamed, so
// a run killed while writing leaves no partial result under the final name.
#include <sys/stat.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/erm.h"

static bool complete(const std::string& p, long long bytes) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && (long long)st.st_size == bytes;
}

int main(int argc, char** argv) {
    erm_params p;
    if (erm_parse_par(argc - 1, (const char* const*)(argv + 1), &p)) {
        fprintf(stderr, "rtm3d: %s\n", erm_last_error(nullptr));
        return 1;
    }
    erm_ctx* ctx = nullptr;
    if (erm_create(&p, 0, &ctx)) {
        fprintf(stderr, "rtm3d: %s\n", erm_last_error(nullptr));
        return 1;
    }
    int nnx, nny, nnz;
    erm_get_dims(ctx, &nnx, &nny, &nnz, nullptr);
    fprintf(stderr, "nnx=%d,nny=%d,nnz=%d\n,gpu_memory_size of one wavefield=%f\n", nnx, nny, nnz,
            (double)nnx * nny * nnz * 4.0 / 1024.0 / 1024.0);
    if (erm_load_model_files(ctx)) {
        fprintf(stderr, "rtm3d: %s\n", erm_last_error(ctx));
        return 1;
    }
    mkdir("result", 0755);
    mkdir("someoutput", 0755);
    const size_t nrec = (size_t)3 * p.lt * p.receiver_num_x * p.receiver_num_y;
    std::vector<float> rec(nrec);
    for (int isz = 0; isz < p.shot_num_z; ++isz)
        for (int isy = 0; isy < p.shot_num_y; ++isy)
            for (int isx = 0; isx < p.shot_num_x; ++isx) {
                if (erm_set_shot(ctx, isx, isy, isz)) {
                    fprintf(stderr, "rtm3d: %s\n", erm_last_error(ctx));
                    return 1;
                }
                int sx = p.shot_start_x + isx * p.shot_interval_x, sy = p.shot_start_y + isy * p.shot_interval_y;
                int sz = p.shot_start_z + isz * p.shot_interval_z;
                const long long bytes = 4LL * p.nx * p.ny * p.nz;
                bool done = true;
                for (const char* img : {"tp", "pp", "ps"}) {
                    std::string f = std::string("result/result-") + img + "-" +
                                    (p.vsp == 0 ? std::to_string(sy) + "-" + std::to_string(sx) : std::to_string(sz));
                    done = done && complete(f, bytes);
                }
                if (done) {
                    fprintf(stderr, "shot isx=%d isy=%d isz=%d: results exist, skipped\n", isx + 1, isy + 1, isz + 1);
                    continue;
                }
                int rc = 0;
                if (p.join_shot) {
                    if ((rc = erm_read_raw(p.infile_shot_name, rec.data(), nrec)) == 0) rc = erm_set_record(ctx, rec.data());
                } else {
                    rc = erm_forward_real(ctx, nullptr);
                }
                if (!rc) rc = erm_forward_smoothed(ctx);
                if (!rc) rc = erm_write_excitation(ctx, "someoutput");
                if (!rc) rc = erm_backward_smoothed(ctx, nullptr);
                if (!rc) rc = erm_write_results(ctx, "result");
                if (rc) {
                    fprintf(stderr, "rtm3d: %s\n", erm_last_error(ctx));
                    return 1;
                }
                fprintf(stderr, "shot isx=%d isy=%d isz=%d done\n", isx + 1, isy + 1, isz + 1);
            }
    erm_destroy(ctx);
    return 0;
}
