#pragma once

#include <cstdint>
#include <vector>

#include "../../include/rovaca_phmm.h"

namespace rvc {

int server_start();
void server_stop();
bool server_running();
// One region through the serving loop (submit + wait), as a worker thread
// would call it.
int served_region(const rvc_region* region, double* out, uint8_t* flags, const DevicePrep* prep = nullptr);
struct DevicePrep;
// Packs a region and queues it on the serving loop (rvc_phmm_submit, with
// device preprocessing when prep is given).
// sites (optional): haplotype -> allele maps to marginalise on the device.
int submit_impl(const rvc_region* region, const DevicePrep* prep, const rvc_sites* sites, rvc_ticket* ticket);
// rvc_phmm_wait, also handing over the request's marginals (when it had sites).
int wait_impl(rvc_ticket ticket, double* out_read_major, uint8_t* out_flags, int64_t timeout_ms,
              rvc_ticket_timing* timing, std::vector<double>* marg);
int region_sync_impl(const rvc_region* region, double* out, uint8_t* flags, const DevicePrep* prep = nullptr);
// The no-allocation one-region call's set-up (its device scratch address and
// stream) at init, and its release at shutdown (phmm_noalloc.cpp).
int noalloc_init();
void noalloc_shutdown();

}  // namespace rvc
