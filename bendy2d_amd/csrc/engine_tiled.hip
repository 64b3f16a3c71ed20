// engine_tiled.hip -- TILED engine (placeholder: falls back to GLOBAL).
#include "engine.h"
namespace bd {
bool tiled_engine_fits(const HostScene &, const Plan &) { return false; }
Engine *make_tiled_engine(hipStream_t s) { return make_global_engine(s); }
}  // namespace bd
