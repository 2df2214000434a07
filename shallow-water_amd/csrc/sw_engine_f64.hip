// The double instantiation of the device engine (Engine<double>: sw_engine.hpp, sw_engine_ops.inc).
#include "sw_engine_ops.inc"

namespace sw {
template struct Engine<double>;
}  // namespace sw
