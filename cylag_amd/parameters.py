"""Solver parameters: defaults, type checks, value checks and the default
time scheme (cylag/core/parameters.pyx:116-352)."""
from .cylag_dico import CYLAG_DEFAULT_PARAMETERS


class Parameters:
    def __init__(self, parameters={}):
        params = dict(CYLAG_DEFAULT_PARAMETERS)
        params.update(parameters)
        self.check_types(CYLAG_DEFAULT_PARAMETERS, params)
        for key, value in params.items():
            setattr(self, key, value)
        # frozen fields start at t = 0 (parameters.pyx:133-135)
        if self.frozen_eulerian_fields == 1:
            self.initial_time = 0.
        self.control_sections = False
        self.ncsec = 0
        # default time scheme (parameters.pyx:179-195)
        if 'time_scheme' not in parameters:
            if self.model == 1 and self.diffusion_model == 0:
                self.time_scheme = 4
            elif self.model == 1 and self.diffusion_model != 0:
                self.time_scheme = 1
            elif self.model == 2:
                self.time_scheme = 5
            elif self.model == 3:
                self.time_scheme = 5

    def check_values(self):
        """Raise ValueError for any out-of-range value (parameters.pyx:197-316)."""
        bad = False
        bad |= self.time_step < 0.
        bad |= self.model < 0 or self.model > 3
        if self.model == 1:
            bad |= self.time_scheme < 1 or self.time_scheme > 4
        elif self.model == 2:
            if self.diffusion_model == 0:
                bad |= self.time_scheme < 1 or self.time_scheme > 5
            else:
                bad |= self.time_scheme != 1 and self.time_scheme != 5
            bad |= self.diffusion_lsm2_option < 1 or self.diffusion_lsm2_option > 3
        elif self.model == 3:
            bad |= self.diffusion_model == 0
            bad |= self.time_scheme != 1 and self.time_scheme != 5
        bad |= self.particle_velocity_init < 0 or self.particle_velocity_init > 1
        bad |= self.rng_method < 1 or self.rng_method > 3
        bad |= self.water_density < 0.
        bad |= self.diffusion_model < 0 or self.diffusion_model > 3
        bad |= self.horizontal_diffusivity < 0. or self.vertical_diffusivity < 0.
        bad |= self.schmidt_number <= 0.
        bad |= self.diffusion_lsm2_option < 1 or self.diffusion_lsm2_option > 3
        bad |= self.diffusion_lsm3_tl_horizontal <= 0. or self.diffusion_lsm3_tl_vertical <= 0.
        bad |= self.lsm3_edge_opt < 1 or self.lsm3_edge_opt > 2
        bad |= self.max_wall_rebounds < 0
        bad |= self.boundary_conditions_type < 1 or self.boundary_conditions_type > 2
        bad |= self.rebound_damping_coef < 0. or self.rebound_damping_coef > 1.
        bad |= self.wall_roughness_angle < 0.
        bad |= self.stranding_probability < 0. or self.stranding_probability > 1.
        bad |= self.listing_printout_period < 0 or self.output_printout_period < 0
        bad |= self.output_file_format not in ['txt', 'vtk', 'all']
        if bad:
            raise ValueError("Failed initialization, check parameters")

    def check_types(self, reference, data, allow_missing_keys=True):
        """Strict type check against the defaults (parameters.pyx:318-352)."""
        errors = {}
        for key, value in data.items():
            if key not in reference:
                errors[key] = "unexpected key"
            elif not isinstance(value, type(reference[key])):
                errors[key] = "expected {}, got {}".format(type(reference[key]).__name__,
                                                           type(value).__name__)
        if not allow_missing_keys:
            for key in reference:
                if key not in data:
                    errors[key] = "missing key"
        if errors:
            print(errors)
            raise TypeError("Failed initialization, type check failed with errors: {}".format(errors))
