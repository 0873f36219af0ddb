// sfqueue_hip.cpp — the reference harnesses include this file textually
// after sfqueue_hip.hpp (throughout_bench.cpp:229-230); every definition of
// the drop-in SFQ already lives in the header.
#include "sfqueue_hip.hpp"
