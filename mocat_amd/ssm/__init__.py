"""mocat.ssm surface (mocat/ssm.py): models, bootstrap, optimal-proposal and ensemble Kalman filters, filtering drivers."""
from .models import StateSpaceModel, TimeHomogenousLinearGaussian, StochasticVolatility
from .filtering import (ParticleFilter, BootstrapFilter, OptimalNonLinearGaussianParticleFilter, EnsembleKalmanFilter,
                        run_particle_filter_for_marginals, run_particle_filters, resample_indices, categorical,
                        run_kalman_filter_for_marginals, run_kalman_smoother_for_marginals,
                        run_kalman_smoother_for_marginals_and_lag_ones, initiate_particles, resample_particles,
                        propagate_particle_filter)
from .backward import backward_simulation, forward_filtering_backward_simulation
