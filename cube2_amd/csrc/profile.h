// Event-based stage profiler (active only when cube2_profile_enable(1)).
#pragma once
#include <string>

#include "common.h"

namespace cube2 {

// Records an event pair around its lifetime on the library stream and the
// stage's algorithmic bytes / flops; resolved by sync_and_flush().
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
};

void profile_flush();
int sync_and_flush();

}  // namespace cube2
