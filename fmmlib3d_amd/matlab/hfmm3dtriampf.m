function [U]=hfmm3dtriampf(iprec,zk,nsource,triaflat,trianorm,source,ifcharge,charge,ifdipole,dipstr,dipvec,ntarget,target,ifpottarg,iffldtarg)
%HFMM3DTRIAMPF Helmholtz triangle FMM with outputs at the targets only, on
% the MI355X (fmm3d-mi355x); the calling sequence and output struct of the
% reference's HFMM3DTRIAMPF (matlab/hfmm3dtriampf.m of fmmlib3d):
%   [U]=HFMM3DTRIAMPF(IPREC,ZK,NSOURCE,TRIAFLAT,TRIANORM,SOURCE,IFCHARGE,CHARGE,IFDIPOLE,DIPSTR,DIPVEC,NTARGET,TARGET)
%   [U]=HFMM3DTRIAMPF(...,NTARGET,TARGET,IFPOTTARG,IFFLDTARG)
if nargin == 13, ifpottarg = 1; iffldtarg = 1; end
[tri, src, q, ds, trg] = fmm3d_mi355x_tri_args(nsource, triaflat, source, ifcharge, charge, ifdipole, dipstr, ntarget, target);
[ier, ~, ~, pottarg, fldtarg] = fmm3d_mi355x_mex('hfmm3dtriatarg', iprec, complex(double(zk)), tri, src, ...
    double(ifcharge), q, double(ifdipole), ds, 0, 0, trg, double(ifpottarg), double(iffldtarg));
if ifpottarg == 1, U.pottarg = pottarg; end
if iffldtarg == 1, U.fldtarg = fldtarg; end
U.ier = ier;
