"""MI355X-native D3Q27 lattice-Boltzmann step of the sim_tcpc path."""
