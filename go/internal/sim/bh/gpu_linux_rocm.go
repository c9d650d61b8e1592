//go:build linux && cgo && rocm

// The gfx950 (MI355X) backend of bh.Solver: the counterpart on Linux with
// ROCm of the Metal walk in gpu_darwin.go. It keeps that file's API
// (GPUAvailable, GPUName, LastGPUSeconds) and runs the whole force step on the
// GPU through the C ABI of libgravitas_hip.so (include/gravitas_hip.h): the
// build (collect, Morton keys, radix sort, octree, moments) as well as the
// walk, so Accelerations hands over the state instead of a built tree.
//
// To install, copy this file to internal/sim/bh/, put the library where the
// cgo lines below point (or adjust them), and narrow gpu_other.go's build
// constraint to `!(darwin && cgo) && !(linux && cgo && rocm)`. In bh.go the
// GPU branch of Accelerations (bh.go:79) becomes
//
//	if b.GPU && GPUAvailable() == nil {
//		b.accelerationsGPU(s, p)
//		return
//	}
//
// placed before b.build, since the GPU builds its own tree. engine.gpu()
// (internal/engine/engine.go:147-155), perf.Calibrate
// (internal/perf/calibrate.go:41-65) and galaxy_set_gpu
// (cmd/libgalaxy/main.go:483-491) need no change: they only call GPUAvailable
// and time Accelerations with GPU on and off.
//
// This file has not been compiled: the environment it was written in has no
// Go toolchain. The C calls it makes are exercised from plain C by
// tests/c_abi_smoke.c with the same pointers and sizes.

package bh

/*
#cgo CFLAGS: -I${SRCDIR}/../../../third_party/gravitas_amd/include
#cgo LDFLAGS: -L${SRCDIR}/../../../third_party/gravitas_amd/gravitas_amd -lgravitas_hip -Wl,-rpath,${SRCDIR}/../../../third_party/gravitas_amd/gravitas_amd
#include <stdlib.h>
#include "gravitas_hip.h"
*/
import "C"

import (
	"errors"
	"sync"
	"unsafe"

	"galaxy/internal/sim"
)

var (
	gpuOnce sync.Once
	gpuErr  error
)

// GPUAvailable reports whether the gfx950 backend can be used, initialising it
// on first use (gpu_darwin.go:132-142). The status codes keep gpu_init's
// meaning: 1 no device, 2 no gfx950 code object.
func GPUAvailable() error {
	gpuOnce.Do(func() {
		var msg [512]C.char
		if C.gx_init(&msg[0], C.int(len(msg))) != C.GX_OK {
			gpuErr = errors.New("hip: " + C.GoString(&msg[0]))
		}
	})
	return gpuErr
}

// GPUName is the name of the device, or "" if the backend is unavailable.
func GPUName() string {
	if GPUAvailable() != nil {
		return ""
	}
	var name [256]C.char
	C.gx_name(&name[0], C.int(len(name)))
	return C.GoString(&name[0])
}

// gpuBuffers holds the Solver's device solver: its arena of device buffers
// and its stream, grown on demand as gpu_darwin.go's buffers are.
type gpuBuffers struct {
	solver      *C.gx_solver
	lastSeconds float64
}

// LastGPUSeconds is the GPU time of the last force step, 0 before the first
// one. The library records it whatever its timing setting.
func (b *Solver) LastGPUSeconds() float64 { return b.gpu.lastSeconds }

func f64(s []float64) *C.double { return (*C.double)(unsafe.Pointer(&s[0])) }

// accelerationsGPU is Accelerations on the GPU: star-star accelerations in
// state order, holes 0, dead bodies untouched, dark matter with DarkMatterOff
// a target of source mass 0 (bh.go:76-102, sim.go:93-98).
func (b *Solver) accelerationsGPU(s *sim.State, p sim.Params) {
	n := s.Len()
	if n == 0 {
		return
	}
	if b.gpu.solver == nil {
		if b.gpu.solver = C.gx_solver_new(); b.gpu.solver == nil {
			panic("gravitas_hip: " + C.GoString(C.gx_error()))
		}
	}
	mono := C.int(0)
	if b.MonopoleOnly {
		mono = 1
	}
	// theta 0 is DefaultTheta in the library too (bh.go:69-74)
	if rc := C.gx_solver_config(b.gpu.solver, C.double(b.Theta), mono, C.GX_WALK_FP32, 0); rc != C.GX_OK {
		panic("gravitas_hip: " + C.GoString(C.gx_error()))
	}
	dm := C.int(0)
	if p.DarkMatterOff {
		dm = 1
	}
	// sim.Kind is a uint8 and Go's bool is one byte (0 or 1): both slices
	// are passed as they are.
	rc := C.gx_accelerations(b.gpu.solver, f64(s.PosX), f64(s.PosY), f64(s.PosZ), f64(s.Mass),
		(*C.uint8_t)(unsafe.Pointer(&s.Kind[0])), (*C.uint8_t)(unsafe.Pointer(&s.Alive[0])),
		C.int64_t(n), C.double(p.Softening), dm, f64(s.AccX), f64(s.AccY), f64(s.AccZ))
	if rc != C.GX_OK {
		panic("gravitas_hip: " + C.GoString(C.gx_error()))
	}
	b.gpu.lastSeconds = float64(C.gx_last_seconds(b.gpu.solver))
}

// walkGPU is not used on this backend: the GPU builds its own tree, and
// Accelerations calls accelerationsGPU before any CPU build.
func (b *Solver) walkGPU(float64) { panic("unreachable: the ROCm backend runs the whole step") }
