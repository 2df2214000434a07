// The float instantiation of the device engine (Engine<float>: sw_engine.hpp, sw_engine_ops.inc).
#include "sw_engine_ops.inc"

namespace sw {
template struct Engine<float>;
}  // namespace sw
