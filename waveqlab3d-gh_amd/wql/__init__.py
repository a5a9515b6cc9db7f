"""Host side of the MI355X WaveQLab3D time-stepping path."""
