//go:build linux && cgo && rocm

// The gfx950 backend of sph.Solver.Forces (sph.go:83-263): the whole hydro
// evaluation (neighbour tree, smoothing lengths, density, pressure forces
// with viscosity, du/dt, div v, Courant step and wake) through the C ABI of
// libgravitas_hip.so, bit-identical to the CPU path.
//
// To install, copy this file and gpu_other_rocm.go to internal/sph/, add a
// field `gpu gpuHydro` to Solver (sph.go:11-19), and start Forces with
//
//	if sv.forcesGPU(s, p, active, step, out) {
//		return
//	}
//
// The backend is used whenever bh.GPUAvailable() reports it usable, as the
// gravity backend is.
//
// This file has not been compiled: the environment it was written in has no
// Go toolchain. The call it makes is exercised from plain C by
// tests/c_abi_smoke.c (a 3x3x3 gas lattice).

package sph

/*
#cgo CFLAGS: -I${SRCDIR}/../../third_party/gravitas_amd/include
#cgo LDFLAGS: -L${SRCDIR}/../../third_party/gravitas_amd/gravitas_amd -lgravitas_hip -Wl,-rpath,${SRCDIR}/../../third_party/gravitas_amd/gravitas_amd
#include "gravitas_hip.h"
*/
import "C"

import (
	"unsafe"

	"galaxy/internal/sim"
	"galaxy/internal/sim/bh"
)

// gpuHydro is the Solver's device hydro: its neighbour tree, buffers and
// stream.
type gpuHydro struct{ h *C.gx_hydro }

func dp(s []float64) *C.double {
	if len(s) == 0 {
		return nil
	}
	return (*C.double)(unsafe.Pointer(&s[0]))
}

// forcesGPU is Forces on the GPU; it returns false (and the caller takes the
// CPU path) when the backend is unavailable.
func (sv *Solver) forcesGPU(s *sim.State, p *sim.Params, active []bool, step []float64, out *sim.HydroOut) bool {
	if bh.GPUAvailable() != nil || s.Len() == 0 {
		return false
	}
	if sv.gpu.h == nil {
		if sv.gpu.h = C.gx_hydro_new(); sv.gpu.h == nil {
			return false
		}
	}
	gas := C.gx_gas_params{gamma: C.double(p.Gas.Gamma), neighbors: C.double(p.Gas.Neighbors),
		alpha: C.double(p.Gas.Alpha)}
	var act *C.uint8_t
	if active != nil {
		act = (*C.uint8_t)(unsafe.Pointer(&active[0])) // Go's bool is one byte, 0 or 1
	}
	rc := C.gx_hydro_forces(sv.gpu.h, dp(s.PosX), dp(s.PosY), dp(s.PosZ), dp(s.VelX), dp(s.VelY), dp(s.VelZ),
		dp(s.Mass), (*C.uint8_t)(unsafe.Pointer(&s.Kind[0])), (*C.uint8_t)(unsafe.Pointer(&s.Alive[0])),
		dp(s.U), dp(s.H), dp(s.Rho), C.int64_t(s.Len()), &gas, C.double(p.Softening), act, dp(step),
		dp(out.AccX), dp(out.AccY), dp(out.AccZ), dp(out.DuDt), dp(out.DivV), dp(out.Dt),
		(*C.uint64_t)(unsafe.Pointer(&out.Wake[0])))
	if rc != C.GX_OK {
		panic("gravitas_hip hydro: " + C.GoString(C.gx_error()))
	}
	return true
}
