"""brainpy.state_amd -- MI355X-native network step of the NEST-style Simulator.

Drop-in names for the reference's ``brainspike`` hot path: ``Simulator``,
``SimulationResult``, ``NodeView``, the connection rules, the neuron models
``iaf_psc_delta`` / ``iaf_psc_exp`` / ``iaf_psc_alpha`` / ``aeif_cond_alpha``,
the devices ``poisson_generator`` / ``spike_recorder`` / ``voltmeter`` /
``multimeter`` and the ``stdp_synapse`` spec. The step runs in hand-written
HIP kernels for gfx950 (``csrc/``) behind the C ABI of ``include/brainspike_hip.h``.
"""
from . import units, init, dist
from .devices import (ac_generator, dc_generator, multimeter, poisson_generator, spike_recorder,
                      step_current_generator, voltmeter)
from .models import aeif_cond_alpha, iaf_psc_alpha, iaf_psc_delta, iaf_psc_exp
from .nodeview import NodeView
from .plasticity import stdp_synapse
from .rules import (all_to_all, explicit_edges, fixed_indegree, fixed_total_number, one_to_one,
                    pairwise_bernoulli)
from .simulator import SimulationResult, Simulator
from .ensemble import Ensemble
from .connections import SynapseCollection
from .shards import Shards

__version__ = '0.1.0'

__all__ = ['Simulator', 'SimulationResult', 'NodeView', 'all_to_all', 'one_to_one',
           'fixed_indegree', 'pairwise_bernoulli', 'fixed_total_number', 'explicit_edges',
           'iaf_psc_delta', 'iaf_psc_exp', 'iaf_psc_alpha', 'aeif_cond_alpha',
           'poisson_generator', 'dc_generator', 'step_current_generator', 'ac_generator', 'spike_recorder',
           'voltmeter',
           'multimeter', 'stdp_synapse',
           'units', 'init', 'dist', 'Ensemble', 'SynapseCollection', 'Shards']
