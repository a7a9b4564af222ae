// GpuPixelSearchByTriplet.h — the reference's entry points, same names and
// signatures (GpuPixelSearchByTriplet.h:12-29 of the reference), served by
// the MI355X library.  Link against cl_forward_amd/libclf.so.
//
// Contract kept from the reference:
//  - output is resized to input.size(); entry i becomes T_i * 100 bytes of
//    Track records whose hits index the event's (x-sorted) hit arrays;
//  - independent_execute sorts every input event by x per sensor IN PLACE
//    (through the const reference, as Tools.cpp:57-63 does), then searches;
//  - gpuPixelSearchByTriplet does no sort and prints nothing;
//  - the return value is 0; a fatal device error prints to stderr and
//    calls exit(-1) (Tools.h:27-34).
// Unlike the reference these functions may be called from several threads
// (calls are serialised on one process-wide context).
#ifndef CLF_GPUPIXELSEARCHBYTRIPLET_H
#define CLF_GPUPIXELSEARCHBYTRIPLET_H

#include <cstdint>
#include <vector>

// GpuPixelSearchByTriplet.cpp:4-22 — standalone path (sorts, verbose).
int independent_execute(const std::vector<std::vector<uint8_t> >& input,
                        std::vector<std::vector<uint8_t> >& output);

// GpuPixelSearchByTriplet.cpp:24-27
void independent_post_execute(const std::vector<std::vector<uint8_t> >& output);

// GpuPixelSearchByTriplet.cpp:29-37 — Gaudi path (no sort, silent).
int gpuPixelSearchByTriplet(const std::vector<const std::vector<uint8_t>*>& input,
                            std::vector<std::vector<uint8_t> >& output);

// GpuPixelSearchByTriplet.cpp:44-65 — common path, chunks of <= 16000 events.
int gpuPixelSearchByTripletInvocation(const std::vector<const std::vector<uint8_t>*>& input,
                                      std::vector<std::vector<uint8_t> >& output);

#endif
