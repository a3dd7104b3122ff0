function [tri, src, q, ds, trg] = fmm3d_mi355x_tri_args(nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ntarget, target)
%FMM3D_MI355X_TRI_ARGS Arrays of the triangle routines in the gateway's
% shapes: the first nsource triangles as 9 x n (the three vertices of
% each), their centroids, the densities (empty when the flag is 0) and the
% first ntarget targets.
tri = reshape(double(triaflat(:, :, 1:nsource)), 9, nsource);
src = double(source(:, 1:nsource));
q = []; ds = []; trg = zeros(3, 0);
if ifcharge ~= 0, q = complex(double(charge(1:nsource))); end
if ifdipole ~= 0, ds = complex(double(dipstr(1:nsource))); end
if ntarget > 0, trg = double(target(:, 1:ntarget)); end
