% Build the MEX gateways against libmesolf (run from this folder in MATLAB).
% libmesolf.so is built by `python -m mesolf_amd.build` one folder up.
here = fileparts(mfilename('fullpath'));
inc = fullfile(here, '..', '..', 'include');
lib = fullfile(here, '..');
mex(['-I' inc], ['-L' lib], '-lmesolf', fullfile(here, 'mex_lfm_convolution_vec.cpp'));
mex(['-I' inc], ['-L' lib], '-lmesolf', fullfile(here, 'recon_phase_space_mex.cpp'));
