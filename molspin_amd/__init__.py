"""MI355X (gfx950) backend for MolSpin's Hilbert-space propagation hot path."""
