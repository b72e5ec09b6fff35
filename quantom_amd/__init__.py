"""MI355X-native implementation of LIMA_MD's per-step path (compound, solvent and
bridge forces, neighbour lists, position-Verlet integration with thermostat)."""
__version__ = "0.1.0"
