// render_hip.cpp - drop-in replacement for the reference's src/render.cpp.
//
// Build the reference's program with this file in place of src/render.cpp and
// link libbhx.so: bh::render (src/render.h:72) then traces on the GPU through
// the C ABI, and every caller (the CLI's stills, exposure probe and animation
// frames, src/main.cpp:906,941,964) is unchanged.  The whole body is the
// header-only adapter of include/bhx_render.hpp.
#include "render.h"

#include "bhx_render.hpp"

namespace bh {

img::Image render(const RenderConfig& cfg, RenderStats& stats) {
    return bhx::render<img::Image>(cfg, stats);
}

}  // namespace bh
