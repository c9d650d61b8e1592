//go:build !(linux && cgo && rocm)

// The stub of gpu_linux_rocm.go for every other build: Forces keeps its CPU
// path. Not compiled here (no Go toolchain in the environment it was
// written in).

package sph

import "galaxy/internal/sim"

type gpuHydro struct{}

func (sv *Solver) forcesGPU(*sim.State, *sim.Params, []bool, []float64, *sim.HydroOut) bool { return false }
