// Drop-in cycleTracking for Quicksilver's own driver.
//
// A driver that keeps the reference's main loop (main.cc:55-66: cycleInit,
// cycleTracking, cycleFinalize on the host, MonteCarlo* mcco) links
// qsg_cycle_tracking.cc in place of main.cc's cycleTracking
// (main.cc:31,138-307) and libqsgpu.so, and adds one call after initMC:
//
//     mcco = initMC(params);
//     qsg_shim_attach(mcco, argc, argv, /*device*/ 0);
//
// Each cycleTracking(mcco) then tracks the processing vaults' particles on
// the GPU and returns the census particles, the balance counters and the
// per-(cell, group) scalar flux into mcco's own structures, so the
// reference's cycleFinalize, its printout, its CORAL2 checks and its next
// cycleInit run unchanged.  One rank and one spatial domain
// (xDom = yDom = zDom = 1).  C++: the argument is the reference's own
// MonteCarlo class.
#ifndef QSG_SHIM_H
#define QSG_SHIM_H

class MonteCarlo;

// Builds the device problem from the same command line and deck as
// getParameters (Parameters.cc:80) and checks that it matches mc (one
// domain, the same cells and energy groups).  Returns 0, or a QSG_ERR_*
// code with the message on stderr.
int qsg_shim_attach(MonteCarlo* mc, int argc, const char* const* argv, int device);
// Releases the device problem.
void qsg_shim_detach();

// main.cc:31 -- the per-cycle tracking of all processing particles.
void cycleTracking(MonteCarlo* monteCarlo);

#endif
