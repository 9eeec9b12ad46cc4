// cnn.h -- host-side handle of the scatter CNN (no HIP types).
#pragma once

#include <memory>

namespace irlamd {
class ScatterCnn;
}

struct irlamd_cnn {
    std::shared_ptr<irlamd::ScatterCnn> net;
};
