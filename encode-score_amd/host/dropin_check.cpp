// Compile check and GPU demo of the drop-in headers: the reference's calls (Kernels::load, Cambi(k, w, h, depth),
// launch(img, pitch, stream), finish()) against libencscore_hip.  Prints the CAMBI score of a synthetic picture.
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
        std::printf("%.17g\n", cambi.finish());
        es_free(img);
        es_stream_destroy(stream);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "encode-score: %s\n", e.what());  // as src/main.cpp:780-783
        return 1;
    }
    return 0;
}
