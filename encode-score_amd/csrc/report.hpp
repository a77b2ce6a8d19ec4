// The run's JSON report in the reference's format (src/main.cpp:693-777), from the scorer's per-segment figures and
// per-frame lists.
#pragma once
#include "encscore.h"

#include <cstdio>
#include <vector>

namespace es {

struct ReportEncode {
    es_segment whole;                  // the segments merged (Accum::merge)
    std::vector<es_segment> segments;  // the segments that have frames, in order
    std::vector<double> vmaf, cambi, cambi_source, psnr_y;  // per frame; vmaf may be shorter (not attached)
};

void write_json(FILE *out, const es_run_info &info, long frames, const std::vector<ReportEncode> &encs);

}  // namespace es
