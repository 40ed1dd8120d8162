function Xguess = recon_theory_phase_space_LFM_with_filter(LFIMG, H, Ht, maxIter, gpuFlag, gpu_id)
% Drop-in replacement for MesoLF's phase-space RL reconstruction
% (src/recon/recon_theory_phase_space_LFM_with_filter.m, same signature and
% output): runs on an AMD Instinct GPU through libmesolf.
%   LFIMG    light-field image [I1, I2], I1 and I2 multiples of Nnum = 15
%   H, Ht    LFM PSFs [S, S, 15, 15, Z]
%   maxIter  RL iterations
%   gpuFlag  must be true (the reference CPU branch cannot run as written)
%   gpu_id   1-based device index
Xguess = recon_phase_space_mex(LFIMG, H, Ht, maxIter, gpuFlag, gpu_id);
end
