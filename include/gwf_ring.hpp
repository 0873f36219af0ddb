// gwf_ring.hpp — drop-in for the reference's G-WFQ implementation header
// (gwf_ring.hpp:1-1380): harnesses include it, alone or after
// gwf_queue.hpp, to get the device API, which lives in gwf_queue.hpp here.
#ifndef GCQ_DROPIN_GWF_RING_HPP
#define GCQ_DROPIN_GWF_RING_HPP
#include "gwf_queue.hpp"
#endif  // GCQ_DROPIN_GWF_RING_HPP
