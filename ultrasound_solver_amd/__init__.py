"""MI355X-native backend for the gpu_iterative Helmholtz path of
Rain-Kotsuzui/ultrasound_solver: hand-written HIP kernels (libhz.so) behind
the reference's solver-backend interface."""
__all__ = ["physics", "operator"]
