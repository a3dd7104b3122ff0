function [src, q, ds, dv, trg] = fmm3d_mi355x_args(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, ntarget, target)
%FMM3D_MI355X_ARGS Arrays of the particle routines in the gateway's shapes:
% the first nsource points and strengths (the reference passes counts
% beside the arrays), empty arrays for terms whose flag is not 1.
src = double(source(:, 1:nsource));
q = []; ds = []; dv = []; trg = zeros(3, 0);
if ifcharge == 1, q = complex(double(charge(1:nsource))); end
if ifdipole == 1
  ds = complex(double(dipstr(1:nsource)));
  dv = double(dipvec(:, 1:nsource));
end
if ntarget > 0, trg = double(target(:, 1:ntarget)); end
