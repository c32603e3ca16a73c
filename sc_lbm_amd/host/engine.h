// engine.h -- internal interface between the Fluid and LBM drop-ins: the
// per-thread recorder of device operators (implemented in lbm.cpp).
#pragma once

class Fluid;
class Geometry;

namespace sclbm_detail {

enum OpKind {
    OP_DENSITY = 0,
    OP_VELOCITIES,
    OP_FEQ_GEOM,
    OP_FEQ_MC,
    OP_REPULSION,
    OP_EQVEL,
    OP_COLLIDE1,
    OP_COLLIDE2,
    OP_VF,
    OP_STREAM1,
    OP_STREAM2
};

// Throws the exception the reference would raise for a C-ABI status.
void raise_status(int rc);
// Runs every recorded operator.
void flush_log();
// Runs the recorded operators that involve f (before f goes away).
void forget(const Fluid* f);
// Records a Fluid method (compute_density, compute_velocities, ...).
void record_fluid_op(int kind, Fluid* f, const Geometry* g);

}  // namespace sclbm_detail
