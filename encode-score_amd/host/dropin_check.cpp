// Compile check and GPU demo of the drop-in headers: the reference's calls (Kernels::load, Cambi(k, w, h, depth),
// launch(img, pitch, stream), finish(); CambiCache(path, x0, y0, w, h, depth), find, add) against libencscore_hip.
// Prints the CAMBI score of a synthetic picture; with a second argument (a file), also keeps it in that file's
// source-CAMBI cache under timestamp 0 and prints what a second cache of the same file finds.
#include "cache.hpp"
#include "cambi.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char **argv)
{
    const int w = 256, h = 144;
    try {
        Kernels k;
        k.load(0);
        es_stream stream = Kernels::stream();
        es_synth_desc d{};
        d.kind = 1;
        d.seed = argc > 1 ? (uint64_t)std::atoll(argv[1]) : 7;
        d.full_w = d.w = w;
        d.full_h = d.h = h;
        d.depth = 10;
        d.grain_q8 = 256;
        d.quant_step = 4;
        d.quant_where = 1;
        void *img = nullptr;
        es_check(es_malloc(&img, (size_t)w * h * 2), "allocating");
        es_check(es_synth_frame(&d, 0, img, 2, stream), "generating");
        Cambi cambi(k, w, h, 10);
        cambi.launch(img, w, stream);
        const double score = cambi.finish();
        std::printf("%.17g\n", score);
        if (argc > 2) {
            {
                CambiCache cache(argv[2], 0, 0, w, h, 10);
                cache.add(0, score);
            }
            CambiCache again(argv[2], 0, 0, w, h, 10);
            const std::optional<double> found = again.find(0);
            std::printf("cached %.17g in %s\n", found ? *found : -1.0, again.file().c_str());
        }
        es_free(img);
        es_stream_destroy(stream);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "encode-score: %s\n", e.what());  // as src/main.cpp:780-783
        return 1;
    }
    return 0;
}
