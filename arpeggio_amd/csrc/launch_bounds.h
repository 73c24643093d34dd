// Workgroup-size bounds of the annealing kernels. Each kernel is compiled with __launch_bounds__ of its constant, and
// the C ABI refuses (ARP_E_ARG, before any HIP call) a launch shape above it, so a mismatch is an argument error on
// the host instead of a launch failure on the device.
#pragma once

#define ARP_ARPEGGIO_MAX_THREADS 512    // arpeggio_fx_kernel: up to 8 anneals (waves) per workgroup
#define ARP_BATCHED_MAX_THREADS 1024    // arpeggio_batched_kernel: up to 16 waves
#define ARP_CHORD_MAX_THREADS 1024      // chord_fx_kernel: up to 16 anneals per workgroup
#define ARP_ARPEGGIO_F32_MAX_THREADS 512
#define ARP_CHORD_F32_MAX_THREADS 256
