// Drop-in for the reference's src/cambi.hpp: the CAMBI banding score of a 10-bit (or 8-bit) luma plane on the GPU.
// launch() queues the whole computation, reading the plane in stream order on the caller's stream and running the
// rest on the instance's own stream; finish() blocks and returns the score clamped to 1000.  Each launch() is
// followed by its finish() before the instance's next launch() (src/cambi.hpp:1-4).
#pragma once
#include "gpu.hpp"

#include <cstdint>

class Cambi {
public:
    // `depth` is the bit depth of the planes it will be given, 8 or 10 (src/cambi.hpp:12).
    Cambi(const Kernels &, int w, int h, int depth) { es_check(es_cambi_create(w, h, depth, &c_), "cambi buffers"); }
    ~Cambi()
    {
        if (c_) es_cambi_destroy(c_);
    }
    Cambi(const Cambi &) = delete;
    Cambi &operator=(const Cambi &) = delete;
    // `img` is a device plane of uint16 samples, `pitch` in pixels (src/cambi.hpp:15-18).
    void launch(const void *img, int pitch, es_stream stream) { es_check(es_cambi_launch(c_, img, pitch, stream), "cambi launch"); }
    double finish()
    {
        double score = 0;
        es_check(es_cambi_finish(c_, &score), "cambi");
        return score;
    }

private:
    es_cambi *c_ = nullptr;
};
