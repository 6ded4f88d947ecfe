//! Drop-in for the reference's `pub mod fluid` (src/lib.rs:4): the smoke
//! solver's compute runs in `libaether_fluid_hip.so` (hand-written HIP
//! kernels for gfx950) behind the C ABI of `include/aether_fluid.h`.
//!
//! * [`ffi`] is the `extern "C"` block, one declaration per entry point of
//!   the header.
//! * [`fluid`] keeps the reference's types and method names
//!   (`FluidSolver3D`, `FluidTelemetryFrame`, `PoissonSolverKind`,
//!   `FluidMedium`), so callers switch by changing the module path.
//! * [`tape`] is the reverse-mode node of a smoke rollout: forward records
//!   the tape on the device, backward returns vector-Jacobian products.

pub mod ffi;
pub mod fluid;
pub mod tape;

pub use fluid::{FluidMedium, FluidSolver3D, FluidTelemetryFrame, MediumInteraction, PoissonSolverKind};
pub use tape::SmokeTape;
