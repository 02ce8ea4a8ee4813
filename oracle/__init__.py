"""FP64 CPU oracle for the aeros transform path (test infrastructure only)."""
