"""CPU FP64 oracle for the FLOWPanel panel-method hot path (test infrastructure only)."""
