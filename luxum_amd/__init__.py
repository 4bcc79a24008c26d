"""MI355X-native Yee FDTD timestep (Luxum drop-in for the src/solvers time loop)."""
