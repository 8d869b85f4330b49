"""Learned corrector: a message-passing network on the mesh graph that predicts the state
a window of solver steps ahead, trained on trajectories of the batched solver, with the
plain Runge-Kutta solver as its fallback."""
