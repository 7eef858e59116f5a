"""radiant_hip -- MI355X (gfx950) backend for Radiant.jl's discrete-ordinates transport solve."""
