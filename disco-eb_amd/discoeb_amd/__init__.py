"""MI355X-native DISCO-EB perturbation hot path (``evolve_perturbations``)."""
