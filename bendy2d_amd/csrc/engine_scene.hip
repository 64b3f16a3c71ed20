// engine_scene.hip -- SCENE engine (placeholder until the workgroup kernel lands).
#include "engine.h"
namespace bd {
bool scene_engine_fits(const HostScene &, const Plan &) { return false; }
Engine *make_scene_engine(hipStream_t s) { return make_global_engine(s); }
bool tiled_engine_fits(const HostScene &, const Plan &) { return false; }
Engine *make_tiled_engine(hipStream_t s) { return make_global_engine(s); }
}  // namespace bd
