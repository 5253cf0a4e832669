// Event-based stage profiler (active only when cube2_profile_enable(1)).
#pragma once
#include <string>

#include "common.h"

namespace cube2 {

// Records an event pair around its lifetime on the library stream (or the
// stream a StageStream names) and the stage's algorithmic bytes / flops;
// resolved by sync_and_flush().
class Stage {
   public:
    Stage(const char* name, double bytes = 0, double flops = 0);
    ~Stage();
    void set_work(double bytes, double flops);

   private:
    bool on_ = false;
    std::string name_;
    double bytes_ = 0, flops_ = 0;
    hipEvent_t a_ = nullptr;
    hipStream_t s_ = nullptr;
};

// Stages opened during its lifetime record on stream s.
class StageStream {
   public:
    explicit StageStream(hipStream_t s) : prev_(ctx().stage_stream) { ctx().stage_stream = s; }
    ~StageStream() { ctx().stage_stream = prev_; }

   private:
    hipStream_t prev_;
};

void profile_flush();
int sync_and_flush();

// Stage times of one call when environment variable env is set (the
// reference's CUBE_PM3_PROFILE / CUBE_PP_TIMING switches): begin returns
// whether profiling is on for the call; end (after the call's work has
// completed) returns " stage=ms ..." and restores the caller's statistics.
bool call_profile_begin(const char* env);
std::string call_profile_end();

}  // namespace cube2
